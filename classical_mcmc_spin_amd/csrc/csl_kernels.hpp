// csl_kernels.hpp — kernel argument blocks and host launchers of
// csl_kernels.hip.
#pragma once

#include <algorithm>

#include "csl_internal.hpp"

namespace csl {

// Persistent batched integration: one workgroup per trajectory.
struct PersistArgs {
    int B;
    const double* init;      // AoS reference order; trajectory b starts at init + b*init_stride
    size_t init_stride;      // 0: all trajectories share one initial state
    const double* t_pulse;   // [B*2]
    const double* amp;       // [B*2]
    const double* tgrid;     // [n_steps+1] step start times (+ final time)
    int n_steps;
    int save_interval;
    int n_rec;
    double dt;
    double* mags;            // [n_rec*B*9] or null
    double* energy;          // [n_rec*B] or null
    double* final_states;    // [B*N*3] AoS or null
    // per-trajectory Hamiltonians: trajectory b reads L.mats + b*mats_stride
    // and L.field + b*field_stride (0 = shared tables)
    size_t mats_stride;
    size_t field_stride;
    // Work split for batches larger than the co-resident workgroups: the
    // launch has `grid` workgroups and workgroup g runs the pieces
    // sched[(g * max_pieces + i) * 5 + {0..4}] = (trajectory, first step, end
    // step, slot to load the state from (-1: the initial state), slot to save
    // the state to (-1: none)); trajectory -1 marks an unused piece.  A piece
    // that loads a slot waits until the piece that saves it has set its flag.
    // sched == null: workgroup g integrates trajectory g from 0 to n_steps.
    const int* sched;
    int max_pieces;
    int grid;
    double* saved;           // [slots][NP][3] device-order states at the seams
    unsigned* flags;         // [slots], zero before the launch
    unsigned* status;        // set to 1 if a wait for a slot gave up
    // Lattices of different sizes in one launch (kPersistMulti): trajectory b
    // runs on lats[lat_of[b]] and starts from init + init_off[b]
    const LatDev* lats;
    const int* lat_of;
    const size_t* init_off;
};

// Partitioned persistent batched integration (csl_partition.hip), for
// lattices whose state does not fit one workgroup's LDS: each trajectory is
// spread over P workgroups ("parts", a contiguous range of c_per cells each,
// in the reference's cell order); a part keeps the RK4 state of its own
// sites in VGPRs and the stage values of its own sites plus a halo of the
// remote neighbours in LDS.  Every stage a part writes the new values of
// the sites other parts read into a global exchange buffer (two parities,
// device site order) and publishes a version number; before a stage it
// waits until every part it exchanges with has published that version.
// The launch holds `groups` groups of P co-resident parts; group g runs
// trajectories g, g + groups, ...
struct PartArgs {
    int B;
    const double* init;        // AoS reference order; trajectory b starts at init + b*init_stride
    size_t init_stride;
    const double* t_pulse;     // [B*2]
    const double* amp;         // [B*2]
    const double* tgrid;       // [n_steps+1]
    int n_steps;
    int save_interval;
    int n_rec;
    double dt;
    size_t mats_stride;        // per-trajectory Hamiltonians (0 = shared tables)
    size_t field_stride;
    int P;                     // parts per trajectory
    int groups;                // trajectory groups (grid = groups * P)
    int xcd_map;               // 1: the parts of a group are placed on one XCD
    int c_per;                 // cells per part
    int cpad;                  // cells per sublattice block of a part, multiple of 64
    int n_loc;                 // n_atoms * cpad local slots (own sites)
    int n_halo_max;            // halo slots per part (LDS slots n_loc ..)
    int n_nbr_max;             // parts a part exchanges with, at most 64
    int dbuf;                  // 1: two LDS buffers of the own stage values (no barrier between
                               // a stage's reads and writes)
    const unsigned* col;       // [P][(zmax+1)/2][n_loc] LDS slots of the bond partners, two 16-bit
                               // slots per word (slot 2j in the low half)
    const int* halo_src;       // [P][n_halo_max] device index of each halo site
    const int* n_halo;         // [P]
    const unsigned* exported;  // [P][n_loc] 1 where another part reads the site
    const int* nbr;            // [P][n_nbr_max] parts exchanged with, -1 unused
    double* exch;              // [groups][2][3][NP]
    unsigned* ver;             // [groups*P] versions published, zero before the launch
    unsigned* status;          // set to 1 when a wait gave up
    double* part_obs;          // [B][n_rec][P][10] observable sums of each part
    double* final_states;      // [B][N*3] AoS or null
    int want_energy;
};
// Sites per thread values of the partitioned kernel and the largest block of each.
bool partitioned_spt_supported(int spt);
int partitioned_block_limit(int spt);
size_t partitioned_lds_bytes(int n_loc, int n_halo_max, int block, bool dbuf);
// Launches the partitioned kernel, or with occupancy != null only stores its
// co-resident workgroups per CU.
hipError_t launch_partitioned(const LatDev& L, const PartArgs& A, int block, int spt, hipStream_t st,
                              int* occupancy);
// mags[(ri*B + b)*9 + j] = inv_n * sum_p part_obs[((b*n_rec + ri)*P + p)*10 + j] (j < 9), energy
// likewise from j = 9 (either output may be null).
hipError_t launch_part_reduce(const double* part_obs, int P, int B, int n_rec, double inv_n, double* mags,
                              double* energy, hipStream_t st);

// Discrete-adjoint gradient run (csl_adjoint.hip).  Four kernels: the
// forward integration (one workgroup per trajectory: records, checkpoints
// and/or the stage inputs of every step), the backward sweep over a segment
// of steps (one workgroup per trajectory: the transposed RK4 stages, storing
// the cotangent u of every site's field at every stage), the parameter
// gradient terms (all CUs: sum_i u_i . dH_i/dtheta at every stage of the
// segment) and the final sums in a fixed order.
struct AdjArgs {
    int B;
    int n_steps;
    int P;                     // parameters (<= 8)
    double dt;
    const double* init;        // AoS reference order; trajectory b at init + b*init_stride
    size_t init_stride;
    const double* fac;         // [n_steps][3][B][2] drive factors or null
    const double* target;      // [n_steps+1][9] target records
    double w[9];               // loss weights of the 9 record components
    const double* dmats;       // [P][n_atoms][zmax][9] dJ/dtheta per (sublattice, slot)
    const double* dfield;      // [P][n_atoms][3]      dB/dtheta per sublattice
    size_t mats_stride;        // per-trajectory Hamiltonians (0 = shared)
    size_t field_stride;
    int n0, n1;                // the launch covers steps [n0, n1)
    int y0;                    // first step held in Y and U
    int seg_len;               // steps held per trajectory in Y and U
    // forward launches
    int ckpt_src;              // >= 0: start from checkpoint ckpt_src (state of step n0); -1: from init
    int record;                // 1: write the records of steps n0 .. n1; 2: only the final one
    int ckpt_interval;         // store x_n every ckpt_interval steps (0: none)
    int n_ckpt;
    double* ckpt;              // [B][n_ckpt][3*NP] checkpoint states (SoA)
    int store_stages;          // store the stage inputs of every step into Y
    double* Y;                 // [B][seg_len][4][3*NP] stage inputs y_1..y_4 of steps n0.. ((x,y) pairs, z plane)
    double* U;                 // [B][seg_len][4][3*NP] cotangents u of the fields at those stages (as Y)
    // backward launches
    double* lam;               // [B][3*NP] adjoint state handed between backward launches (SoA)
    int first;                 // the launch holds the last step: lambda starts from dl/dx_N
    double* rec;               // [n_steps+1][B][9] forward records
    double* loss_n;            // [B][n_steps+1] loss term of every record
    double* part;              // [B][n_steps][4][P] parameter-gradient term of every stage
    double* grad_init;         // [B][N*3] dL/dx_0 (AoS), written by the launch with n0 == 0, or null
    double* loss;              // [B] (final kernel)
    double* grad;              // [B][P] (final kernel)
};

size_t adjoint_lds_bytes(const LatDev& L, int block);
int adjoint_sites_per_thread(const LatDev& L);
hipError_t launch_adjoint_forward(const LatDev& L, const AdjArgs& A, hipStream_t st);
// records of steps A.n0 .. A.n1-1 from the stored stage inputs (A.Y)
hipError_t launch_adjoint_records(const LatDev& L, const AdjArgs& A, hipStream_t st);
hipError_t launch_adjoint_backward(const LatDev& L, const AdjArgs& A, hipStream_t st);
// blocks_per_traj: workgroups per trajectory (each loops over its share of
// the segment's stage slots)
hipError_t launch_adjoint_params(const LatDev& L, const AdjArgs& A, int blocks_per_traj, hipStream_t st);
hipError_t launch_adjoint_final(const AdjArgs& A, hipStream_t st);

// Dormand-Prince 5(4) integration (csl_dopri5.hip): one workgroup per trajectory.
struct Dopri5Args {
    int B;
    const double* init;        // AoS reference order; trajectory b at init + b*init_stride
    size_t init_stride;
    const double* t_pulse;     // [B*2]
    const double* amp;         // [B*2]
    int control;               // 0 fixed grid steps, 1 integrate_adaptive, 2 integrate_const + controller
    const double* tgrid;       // control 0: observation grid [n_grid]
    const double* segs;        // control 2: [B][max_seg] segments (t0, t1), or null for [T0, T1]
    const int* n_segs;         // [B] segments per trajectory (>= 1)
    int max_seg;
    int n_grid;
    double T0, T1, dt;
    double abs_tol, rel_tol;
    int save_interval;
    int max_rec;
    long long max_attempts;    // attempts per trajectory before giving up (status 2)
    double* times;             // [B][max_rec]
    double* mags;              // [B][max_rec][9] or null
    double* energy;            // [B][max_rec] or null
    double* final_states;      // [B][N*3] AoS or null
    long long* counters;       // [B][4]: observer saves, accepted, rejected, status
    size_t mats_stride;        // per-trajectory Hamiltonians (0 = shared)
    size_t field_stride;
};

// status codes in Dopri5Args::counters[b*4 + 3]
constexpr long long kDopri5Ok = 0;
constexpr long long kDopri5StepFailed = 1;   // 500 rejections in a row
constexpr long long kDopri5AttemptCap = 2;   // max_attempts reached
constexpr long long kDopri5BadStep = 3;      // step size no longer advances t or is not finite

// Sites per thread and workgroup size of the Dormand-Prince kernel for a
// lattice (0 when the lattice does not fit one workgroup).
int dopri5_sites_per_thread(const LatDev& L);
int dopri5_block(const LatDev& L);
size_t dopri5_lds_bytes(const LatDev& L, int block);
hipError_t launch_dopri5(const LatDev& L, const Dopri5Args& A, hipStream_t st);

// zero-temperature relaxation (csl_quench.hip)
hipError_t launch_quench_class(const LatDev& L, double* x, const int* color, int c, const int* frozen, double len,
                               int B, hipStream_t st);
hipError_t launch_quench_check(const double* e, double* e_prev, int* frozen, int* sweeps, int B, int sweep,
                               double rel_tol, hipStream_t st);

// structure factor (csl_dssf.hip)
hipError_t launch_sq_fill(const LatDev& L, const double* x, int B, double2* out, hipStream_t st);
hipError_t launch_sq_combine(const double2* F, int B, int n_atoms, int L1, int L2, int L3, const double* pos,
                             double2* sqt, hipStream_t st);
hipError_t launch_sq_power(const double2* sw, int T, int B, size_t Q, double scale, double* out, hipStream_t st);


// One RK4 stage over all sites of B trajectories (SoA [B][3][NP] buffers).
struct StageArgs {
    int B;
    int stage;               // 0..3
    double* x;               // state at step start (updated in stage 3)
    double* acc;             // RK4 accumulator
    const double* ycur;      // stage input (must be x in stage 0)
    double* ynext;           // next stage input (unused in stage 3)
    const double* fac;       // [B*2] drive factors of this stage, or null (no drive)
    double dt;
    int rec;                 // stage 0 of a recorded step: write block partials
    double* partials;        // [B][nblk][10]
};

// Largest workgroup a persistent instantiation accepts.  Smaller bounds give
// each wave a larger VGPR budget (512 / waves-per-SIMD): FP64 state and
// accumulator cost 12 VGPRs per owned site, so threads that own more sites
// run in smaller workgroups.
template <int SPT>
constexpr int persistent_max_block() {
    return SPT <= 3 ? 1024 : (SPT <= 4 ? 768 : 512);
}

// Launcher of the persistent kernel for one sites-per-thread value; each
// specialisation is defined in its own object (csl_persistent.hip built with
// -DCSL_SPT=n).
// Launch-shape flags of the persistent kernel.
constexpr int kPersistDoubleBuffer = 1;  // two LDS state buffers, one barrier per stage
constexpr int kPersistUnrollStages = 2;  // the four RK stages specialised at compile time
constexpr int kPersistMulti = 4;         // per-trajectory lattices (PersistArgs::lats)

// Compile-time specialisation the launcher picks for a persistent launch
// shape (reported by csl_persistent_candidates; the launcher applies the same
// rule): FIELD_ONLY for double-buffered, stage-unrolled shapes of lattices
// with translation-invariant tables, per-sublattice fields, no onsite matrix
// and full coordination; ATOM_MAJOR in addition when the block tiles the
// sublattices (npad = waves_per_atom * 64 * SPT, block = n_atoms *
// waves_per_atom * 64).
constexpr int kVariantFieldOnly = 1;
constexpr int kVariantAtomMajor = 2;
constexpr int kVariantPartitioned = 4;  // reported for launches of the partitioned kernel
inline int zmax_bucket(int zmax) { return zmax <= 3 ? 3 : (zmax <= 6 ? 6 : 12); }

// Atom-major tiles: the npad / 64 wave-sized units of each sublattice are
// dealt to its nwaves / n_atoms waves (at most spt per wave) in order, every
// wave taking the floor share and the remainder going one unit at a time to
// the wave of the sublattice whose SIMD (wave % 4: waves are dealt to the 4
// SIMDs in turn) carries the fewest units so far (ties: the lower wave), so
// that the busiest SIMD, which sets the time of a stage, carries as few as
// possible.  Wave `wave` owns units [start, start + units) of its
// sublattice.  With npad = waves per atom * 64 * spt every wave takes spt.
__host__ __device__ inline void atom_major_tile(int nwaves, int n_atoms, int U, int wave, int& start, int& units) {
    const int wpa = nwaves / n_atoms;
    const int q = U / wpa, r = U % wpa;
    const int a_w = wave / wpa;
    int load[4] = {0, 0, 0, 0};
    unsigned extra = 0;  // waves of sublattice a_w given q + 1 units
    for (int a = 0; a <= a_w; ++a) {
        for (int i = 0; i < wpa; ++i) load[(a * wpa + i) & 3] += q;
        unsigned ex = 0;
        for (int e = 0; e < r; ++e) {
            int best = -1;
            for (int i = 0; i < wpa; ++i)
                if (!((ex >> i) & 1u) && (best < 0 || load[(a * wpa + i) & 3] < load[(a * wpa + best) & 3])) best = i;
            ex |= 1u << best;
            ++load[(a * wpa + best) & 3];
        }
        extra = ex;
    }
    const int i_w = wave - a_w * wpa;
    start = 0;
    for (int i = 0; i < i_w; ++i) start += q + int((extra >> i) & 1u);
    units = q + int((extra >> i_w) & 1u);
}

inline int persistent_variant(const LatDev& L, int block, int spt, int flags) {
    if (flags & kPersistMulti) return 0;
    const bool dbuf = flags & kPersistDoubleBuffer, unroll = flags & kPersistUnrollStages;
    if (!(unroll && dbuf && L.uniform && L.site_uniform && !L.has_onsite && L.full_coord &&
          L.zmax == zmax_bucket(L.zmax)))
        return 0;
    // atom-major: whole waves per sublattice, each with 1..spt units
    const int U = L.npad / kWave, wpa = block / (L.n_atoms * kWave);
    const bool am = L.npad % kWave == 0 && block % (L.n_atoms * kWave) == 0 && wpa * spt >= U && wpa <= U &&
                    wpa <= 32;
    return kVariantFieldOnly | (am ? kVariantAtomMajor : 0);
}

// Busiest SIMD's units of an atom-major launch shape
inline int atom_major_max_load(int nwaves, int n_atoms, int U) {
    int load[4] = {0, 0, 0, 0};
    for (int w = 0; w < nwaves; ++w) {
        int s0 = 0, u = 0;
        atom_major_tile(nwaves, n_atoms, U, w, s0, u);
        load[w & 3] += u;
    }
    return std::max(std::max(load[0], load[1]), std::max(load[2], load[3]));
}

// Launches the persistent kernel, or with occupancy != null only stores its
// co-resident workgroups per CU for this shape.
template <int SPT>
hipError_t launch_persistent_spt(const LatDev& L, const PersistArgs& A, int block, int flags, size_t lds,
                                 hipStream_t st, int* occupancy);
template <> hipError_t launch_persistent_spt<1>(const LatDev&, const PersistArgs&, int, int, size_t, hipStream_t, int*);
template <> hipError_t launch_persistent_spt<2>(const LatDev&, const PersistArgs&, int, int, size_t, hipStream_t, int*);
template <> hipError_t launch_persistent_spt<3>(const LatDev&, const PersistArgs&, int, int, size_t, hipStream_t, int*);
template <> hipError_t launch_persistent_spt<4>(const LatDev&, const PersistArgs&, int, int, size_t, hipStream_t, int*);
template <> hipError_t launch_persistent_spt<6>(const LatDev&, const PersistArgs&, int, int, size_t, hipStream_t, int*);

size_t persistent_lds_bytes(const LatDev& L, int block, bool dbuf);
bool persistent_spt_supported(int spt);
int persistent_block_limit(int spt);
// Co-resident workgroups per CU of the persistent kernel for this launch
// shape (0 when it cannot be determined).
int persistent_occupancy(const LatDev& L, int block, int spt, int flags);
hipError_t launch_persistent(const LatDev& L, const PersistArgs& A, int block, int spt, int flags,
                             hipStream_t st);
int stage_blocks(const LatDev& L);
hipError_t launch_stage(const LatDev& L, const StageArgs& A, hipStream_t st);
hipError_t launch_observe(const LatDev& L, const double* y, double* partials, int B,
                          hipStream_t st);
hipError_t launch_reduce_obs(const double* partials, int nblk, int B, int ri, double inv_n,
                             double* mags, double* energy, hipStream_t st);
hipError_t launch_eval(const LatDev& L, const double* y, double* H, double* dsdt, double f1,
                       double f2, int drive_on, hipStream_t st);
hipError_t launch_perm_to_soa(const LatDev& L, const double* aos, size_t aos_stride, double* soa,
                              int B, hipStream_t st);
hipError_t launch_perm_to_aos(const LatDev& L, const double* soa, double* aos, int B,
                              hipStream_t st);
hipError_t launch_normalize(double* s, int N, double len, hipStream_t st);
hipError_t launch_drive_table(const double* tgrid, int n_steps, double dt, int B,
                              const double* t_pulse, const double* amp, double width, double freq,
                              double* fac, hipStream_t st);

// One element-wise combination g*x + sum_i w[i] * K[idx[i]] of the generic
// fixed-step stage engine (the methods of the reference GPU stepper other
// than rk4): K holds the stage derivatives, n doubles each.
struct RkCombine {
    int m;
    int idx[13];
    double g;
    double w[13];
};
// p[i] = v for i < n (clears flags and counters without a memset node).
hipError_t launch_fill_u32(unsigned* p, unsigned v, size_t n, hipStream_t st);
// out[i] = g*x[i] + sum_j w[j] * K[idx[j]*n + i] for i < n; out may be x.
hipError_t launch_rk_combine(double* out, const double* x, const double* K, size_t n, const RkCombine& c,
                             hipStream_t st);

}  // namespace csl
