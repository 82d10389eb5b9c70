// Device-side layout of the engine: per-replica parameters, per-replica
// dynamic state and the flat device buffers shared by all replicas.
//
// Replicas are concatenated: replica r owns particles [p0, p0 + n), bonds
// [bond0, bond0 + nbond) and so on; every bead index stored on the device is
// already global, so kernels never add offsets to topology entries.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tis {

constexpr int kNT = 9;            // parameter types
constexpr int kMaxCells = 8192;   // cells per replica (LDS counting sort)
constexpr int kBlock = 256;
constexpr int kExcl = 8;          // bonded and 1-3 partners per particle
constexpr int kMaxSlots = 32;     // words of a replica's largest-displacement maximum
constexpr int kSlotStride = 16;   // one 128-byte line per word (unsigned long long units)
constexpr int kCloseStride = 32;  // one 128-byte line per replica's close counter (int units)

// Static parameters of one replica (read-only on the device).
struct RepParams {
    int p0, n, nbeads, kh;        // particle range, RNA beads, recentring bead (global)
    double L[3], hL[3];
    double T, lB, h, vcap;        // vcap = 300 T (speed cap, fast.cpp:604-620)
    double rsr2, rlr2, skin;
    int M[3];
    double lc[3];
    int sr_h;                     // cells searched either side for a short-range-only list (1 when
                                  // every cell is at least r_sr wide, else 2)
    int cell0;                    // offset into the cell-start array (ncell + 1 entries)
    long long pool0, pool_cap;    // neighbour-pool region: n slots of cap_sr + cap_lr entries
    int cap_sr, cap_lr;           // SR and LR entries per particle slot
    int bond0, nbond, ang0, nang;
    int st30, nst3, st20, nst2;
    int hb0, nhb, site0, nsite, inc0, ninc;  // HB candidates, sites, site incidences
    int lr_every;
    double mass[kNT], K1[kNT], K2[kNT], sigma[kNT];
    // pair tables, index ta * 9 + tb
    double d2lj[kNT * kNT], dlj[kNT * kNT], flj[kNT * kNT], elj[kNT * kNT], d2min[kNT * kNT],
        qlb[kNT * kNT];
};

// Dynamic state of one replica (read-write on the device).
struct RepState {
    long long klok;               // step about to be executed
    double progress;              // PROGRESS, fast.cpp:1768-1775
    int need_rebuild;
    int lr_due;
    long long n_rebuilds;
    int n_act;                    // live HB candidates after the last rebuild
    int n_asite;                  // sites with at least one live candidate
    int n_ainc;                   // entries of the per-site live lists
    int overflow;                 // bit 1 pool, 2 HB caps, 4 stream, 8 cell grid
    long long stream_cursor;
    long long hb_draws;           // Philox draws of the current step's HB resolution
    int hb_iters;
    // HB rebuild tables: two buffers; `tab` is the one the resolution reads.  A
    // rebuild inside a step writes the other one (counts in nx_*) and sets
    // tab_pending; the step's close then switches (tab_fresh for one step after).
    int tab, tab_pending, tab_fresh;
    int nx_act, nx_asite, nx_ainc;
    int max_sr, max_lr;           // largest per-particle SR / LR counts of the last build
    int crop_off;                 // set (and kept) once a cell sort met a coordinate outside the box
                                  // after one wrap: list searches then take whole cells
    unsigned rng_epoch;           // Philox key word: a restart of klok on the same context moves to a
                                  // new epoch, so no draw of an earlier run repeats
    double e[12];                 // energies of the last force evaluation
};
enum { EI_E2_HB, EI_E3_HB, EI_E2_ST, EI_E3_ST, EI_WCA, EI_COUL_SR, EI_COUL_LR, EI_BOND, EI_ANGLE };

// Caps of the HB resolution working set held in LDS (per replica).
constexpr int kActMax = 3072;     // live candidates
constexpr int kASiteMax = 3072;   // sites with live candidates
constexpr int kAIncMax = 8192;    // incidences of live candidates

// Phase trace of the HB resolution: a ring of kHbTraceSteps steps (slot klok %
// kHbTraceSteps); per (slot, replica, group) kHbTraceWords words: the real-time
// clock (100 MHz) at the group's start and at the end of each phase (tables in
// LDS, excess ranked, components joined, roots ranked, resolution, outcome and
// forces, end), the serial depth, then n_over | n_comp << 16 | n_live << 32.
constexpr int kHbTraceSteps = 64;
constexpr int kHbTraceWords = 10;

struct Dev {
    int n_rep;
    int n_tot;
    int precision;
    int rng;
    uint64_t seed;
    double noise_scale;
    int pair_lanes;               // lanes per particle in the pair gather
    int hb_groups;                // workgroups per replica in the HB resolution
    int hb_lag;                   // HB tables rebuilt beside the resolution, used from the next step
    long long tab_hb, tab_site, tab_inc;   // elements per buffer of the HB rebuild tables
    int st_slices;                // stack workgroups per replica in a step
    RepParams *rp;
    RepState *rs;
    // particles (global index).  The state and the multi-body forces have two
    // buffers each: a step reads pos / old / vel / f_mb and writes the new state
    // to the *_nx buffers (clearing f_mb_nx for the next step); the host then
    // swaps the pointers
    double4 *pos, *old, *vel;
    double4 *pos_nx, *old_nx, *vel_nx;
    double4 *f_sr, *f_lr, *f_mb, *f_mb_nx;   // f_mb: RNA beads only (ions' entries stay zero)
    double4 *nz;                  // random force of the step (Philox draws; k_noise)
    int *close_cnt;               // k_integrate: workgroups done per replica (kCloseStride apart)
    unsigned long long *maxd2;    // bits of max |x - x_old|^2 of the current step: kMaxSlots words
                                  // per replica, word q of replica r at (r kMaxSlots + q) kSlotStride
    double4 *jv, *jf;             // force matching: position tangent in, force tangent out
    int *ptype;                   // type | cls << 4 | charged << 8
    int *prep;                    // replica of each particle
    // cells and lists
    int *cell_of, *cell_rank, *cell_start, *sorted;
    int *cell_cnt;                // particles per cell of a rebuild in progress (k_bin_count; cleared by k_bin_scan)
    double4 *cpos;                // positions in cell order (current: k_integrate updates them),
                                  // .w = particle index (low word) and type word (high word)
    int *crank;                   // cell-order index of each particle (last rebuild)
    int *excl;                    // kExcl bonded / 1-3 partners per particle (global, -1 none)
    unsigned char *bsub;          // per bonded term: subtract its pair's WCA in the step path too
    int *nb_start, *nb_nsr, *nb_nlr;
    int4 *nb_meta;                // by cell-order index: pool offset of the particle's slot / 16 (slots
                                  // are multiples of 16 entries), SR and LR counts, replica (the pair
                                  // gather loads it beside cpos)
    int *pool;                    // neighbour entries: cell-order indices (~q for bonded / 1-3)
    // topology
    int2 *bond_ij;
    double2 *bond_kr;
    int4 *ang_ijk;
    double2 *ang_kt;
    int *st3_b;                   // 6 per stack
    double *st3_g;                // 6 per stack
    double *st3_E;
    int2 *st3_dis;                // secondary stacks switched off (replica-local index, -1 none)
    int *st2_b;                   // 8 per stack (7 used)
    int *mb_terms;                // k_multibody: each replica's terms by type, sorted by lowest bead
    int *mb_off;                  // k_multibody: 4 offsets into mb_terms per (replica, slice boundary)
    int *mb_win;                  // k_multibody: bead window (first, end) per (replica, slice)
    double2 *st2_rE;
    int *hb_b;                    // 6 per HB
    double *hb_g;                 // 6 per HB
    double *hb_E;
    int *hb_dode;
    int *hb_inc_off, *hb_inc;     // sites of each HB (global site ids), CSR over global HB ids
    int *site_off, *site_hb;      // HBs of each site (global HB ids), CSR over global site ids
    int *site_val;
    int *inc_rank;                // position of each HB incidence within its site's static list
    int *site_cnt;                // live incidences per site (rebuild scratch)
    int *chunk_cnt;               // live candidates per status chunk, kMaxHbChunks per replica
    // HB dynamic (rebuild products)
    signed char *hb_status;       // -1 / 0 per candidate
    int *hb_iter_part;            // per replica and resolution group: serial depth of its wavefronts
    double *hb_e_part;            // per replica and resolution group: its E2, E3 HB sums
    signed char *hb_out;          // outcome of the last HB evaluation: -1 not live-listed, 0 gated, 1 kept, 2 off
    unsigned long long *hb_trace; // phase clock of the resolution (tis_trace_hb), nullptr when off
    // per replica and live index (r * kActMax + a), written by k_hbenergy: energy,
    // gate flag and the six bead forces (18 values of the force precision)
    double *hb_en;
    unsigned char *hb_gate;
    void *hb_frc;
    int *act_hb;                  // per replica region [hb0, hb0 + nhb): live candidate ids
    int *act_tmp;                 // rebuild scratch: each status chunk's live ids, packed
    unsigned char *act_flag;      // region [hb0, hb0 + nhb): gate / resolution outcome
    int *asite_val;               // region [site0, ...): valence of each active site
    int *asite_off;               // region [site0 + rep, ...]: CSR offsets (n_asite + 1)
    int *asite_list;              // region [inc0, ...): live indices per active site, static order
    int *ahb_inc_off;             // region [hb0 + rep, ...]: CSR offsets per live candidate
    int *ahb_inc;                 // region [inc0, ...): active-site index of each incidence
    int4 *act_rec;                // region [2 hb0, ...): candidate id and its six beads, per live index
    // output / scratch
    double *st_energy;            // per tertiary stack
    double *st_part;              // per replica and stack slice: E2, E3 stack sums of the step
    double *maxwell;              // 3000 per replica (stream mode)
    int32_t *stream;              // raw glibc values, region per replica
    long long *stream_base, *stream_len;
    // replica exchange
    double *remd_e;               // 4 per replica: non-HB energy of own configuration, its HB energy,
                                  // non-HB energy of the partner's configuration, spare
    double *remd_d;               // per neighbour pair: exponent D of the last attempt
    long long *remd_stats;        // 2 per neighbour pair: attempts, accepts
    int *remd_ok;                 // per neighbour pair: exchangeable
    int *remd_state;              // per neighbour pair: 0 idle, 1 swapped, 2 to restore
    int *walker;                  // trajectory id held by each replica
    // equilibrium observables (accumulated per sample; layout in observe.h)
    double *obs_rep;
    double4 *obs_p;
    double *obs_hb, *obs_st3;
    int *obs_cnt;
    double *obs_keep;             // a step's energies and stack energies kept across a sample
    signed char *obs_keep_hb;     // and its HB outcomes
};

}  // namespace tis
