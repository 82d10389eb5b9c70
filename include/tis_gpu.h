/*
 * tis_gpu.h - C ABI of the MI355X TIS-RNA + explicit-ion Langevin engine.
 *
 * The reference (src/16SCD_Mg_Cl_K_COPY_fast.cpp, "fast.cpp" below) has no
 * plugin or FFI interface: its hot path is four void() calls on process-global
 * arrays, made once per Langevin step from main():
 *     deterministic_forces_MPI(); full_forces(); move_rigid_units(); check_shifts();
 * (fast.cpp:6380-6386), plus populate_lists() at start-up (fast.cpp:6237).
 * Each entry point below names the reference routine(s) it replaces.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative TIS_E* code; none aborts.
 *     tis_last_error() gives a human-readable message for the calling thread.
 *   - Host buffers passed in are copied; host buffers passed out are filled.
 *   - Particle order is the reference's, shifted to 0-based: reference particle i
 *     (1..NT_atom) is index i-1 here.  Beads first (S(n)=3n-3, B(n)=3n-2 and
 *     P(n)=3n-4 in 0-based terms for 16SCD, which has no 5' phosphate), then Mg,
 *     then Cl, then K.  Vectors are xyz-interleaved double[3*N].
 *   - One tis_ctx per host thread; all device work runs on the context's stream.
 */
#ifndef TIS_GPU_H
#define TIS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
enum {
    TIS_OK = 0,
    TIS_EINVAL = -1,    /* bad argument */
    TIS_ENOMEM = -2,    /* host or device allocation failed */
    TIS_EIO = -3,       /* input file missing or malformed */
    TIS_EDEVICE = -4,   /* HIP runtime error */
    TIS_EOVERFLOW = -5, /* a device list or random stream ran out of room */
    TIS_ESTATE = -6,    /* call out of order (e.g. step before set_state) */
    TIS_ENONFINITE = -7 /* a replica's coordinates became NaN or infinite (the run stops) */
};
const char *tis_last_error(void);

/* ------------------------------------------------------------ the system */
/* Number of parameter types (maxi_key 1..9: P S A G C U Mg Cl K, fast.cpp:2883-2945). */
#define TIS_NTYPES 9

/*
 * One replica: topology + force-field tables + state, all 0-based.
 * Filled by tis_build_16scd() (the reference's set-up, fast.cpp:2883-6031) or
 * by the caller.  Arrays are owned by whoever allocated the struct.
 */
typedef struct tis_system {
    /* sizes */
    int n_particles;   /* NT_atom */
    int n_beads;       /* NTP_atom: RNA beads come first */
    int n_units;       /* NBP_atom: units 1..n alternate nucleoside (odd) / phosphate (even) */
    int n_bonds, n_angles;
    int n_st3, n_st2;  /* tertiary (s3_N) and secondary (st_N - s3_N) stacks */
    int n_hb;          /* hb_N candidates */
    int n_hb_sites;    /* HB_NT_atom */
    int n_hb_incid;    /* total site incidences (sum over HBs of sites listed) */
    int n_mg, n_cl, n_k;
    /* box and thermodynamics */
    double box[3];     /* side_X, side_Y, side_Z */
    double T;          /* kcal/mol */
    double lB;         /* Bjerrum length, fast.cpp:6109 */
    double h;          /* time step in tau, fast.cpp:28 */
    int recentre;      /* 0-based index of the bead kept at the box centre (ref. bead 500) */
    /* per type t = maxi_key-1 */
    double type_mass[TIS_NTYPES], type_R[TIS_NTYPES], type_eps[TIS_NTYPES], type_q[TIS_NTYPES];
    /* per particle */
    int *type;         /* maxi_key - 1, 0..8 */
    int *cls;          /* part_key: 0 P, 1 S, 2 B, 3 Mg, 4 Cl, 5 K */
    double *pos, *pos_old, *vel;   /* 3*N each */
    /* units: unit u (0-based) is reference unit u+1 */
    int *unit_bead0;   /* atom_key[u][0]: S for nucleosides, P for phosphates */
    int *unit_bead1;   /* atom_key[u][1]: B for nucleosides, -1 for phosphates */
    char *unit_kind;   /* amino_key: 'A','C','G','U' or 'P' */
    /* backbone (polymer_constraints, fast.cpp:2354-2520), in the reference's term order */
    int *bond_ij;      /* 2 per bond */
    double *bond_kr;   /* k, r0 per bond */
    int *angle_ijk;    /* 3 per angle, vertex in the middle */
    double *angle_kt;  /* k, theta0 per angle */
    /* tertiary stacks: beads S_i B_i P_i+1 S_j B_j P_j+1, geometry r th1 th2 psi psi1 psi2, E */
    int *st3_beads;    /* 6 per stack */
    double *st3_geo;   /* 6 per stack */
    double *st3_E;
    int *st3_face;     /* ATOM_ST[k][0..1]: 1-based unit index of each face */
    int *st3_disable;  /* 2 per stack: secondary stack switched off when formed, or -1 */
    /* secondary stacks: beads P_n S_n B_n P_n+1 S_n+1 B_n+1 P_n+2, r0, E */
    int *st2_beads;    /* 7 per stack */
    double *st2_r0, *st2_E;
    int *st_excess;    /* ST_EXCESS per unit face (n_units) */
    /* hydrogen-bond candidates: beads k10 k11 k12 k20 k21 k22, geometry r th1 th2 psi psi1 psi2 */
    int *hb_beads;     /* 6 per HB */
    double *hb_geo;    /* 6 per HB */
    double *hb_E;
    int *hb_dode;      /* 1 = tertiary class (E3_HB), 0 = secondary (E2_HB) */
    int *hb_site_off;  /* n_hb+1: ATOM_HB as CSR, sites per HB with multiplicity */
    int *hb_site_idx;
    int *site_hb_off;  /* n_hb_sites+1: HB_ATOM as CSR, HBs per site in insertion order */
    int *site_hb_idx;
    int *site_valence; /* VALENCE */
} tis_system;

/* Arguments of the reference command line (src/run_sample.sh, fast.cpp:6080-6103). */
typedef struct tis_build_args {
    double ss_D;   /* -D */
    double hs_D;   /* -E */
    double st_D;   /* -d */
    double T;      /* -T, kcal/mol */
    double c_mg;   /* -c, mol/L MgCl2 */
    double c_k;    /* -f, mol/L KCl */
    int seed;      /* -s */
    double box;    /* -L (cubic side, A); 350 is the reference's hard-coded value */
} tis_build_args;

/*
 * Builds the 16SCD system from the reference's input files in data_dir
 * (16SCD.pdb, 16SCD_16SCD_unprocessed_bonds.dat, ..._stacks.dat,
 * MAXIKEY_Mg_Cl_K.dat).  Replaces read_maxi_key, list_crowder_types,
 * read_PDB, initialize_HB_atoms, initialize_unprocessed_bonds,
 * initialize_unprocessed_stacks, set_box and add_crowder, called in the
 * order of fast.cpp:6105-6177, including the glibc rand() stream seeded with
 * srand((unsigned)(T*1000*seed)).  *rand_state_out (may be NULL) receives the
 * number of rand() draws consumed, so a caller can continue the same stream.
 */
int tis_build_16scd(const char *data_dir, const tis_build_args *args,
                    tis_system **out, long long *rand_draws_out);

/* Input files of any chain in the reference's formats. */
typedef struct tis_chain_files {
    const char *pdb;            /* atoms as read_PDB reads them (tokens 1-9 of ATOM lines) */
    const char *bonds;          /* HB candidate lines, <name>_<name>_unprocessed_bonds.dat */
    const char *stacks;         /* tertiary stack pairs, <name>_<name>_unprocessed_stacks.dat */
    const char *maxikey;        /* MAXIKEY_Mg_Cl_K.dat */
    const int *double_face_res; /* PDB residue numbers whose 5' stacking face has ST_EXCESS
                                   lowered by 2 (may be NULL) */
    int n_double_face;
    int recentre_bead;          /* 1-based bead kept at the box centre; <= 0 means 500 */
} tis_chain_files;

/*
 * Builds any chain from files in the reference's formats.  The reference's
 * set-up is 16SCD-only: residue numbers map to nucleotides through hard-coded
 * offsets (fast.cpp:3783-3795, 4295-4305) and 15 stack faces are hard-coded
 * exceptions (fast.cpp:4458-4491).  Here residue numbers map to nucleotides in
 * the order the PDB lists them (a number used twice, or a file naming a residue
 * the PDB lacks, is TIS_EIO) and the exceptions are an argument.  Everything
 * else, including the rand() stream, is tis_build_16scd's; that function is
 * this one on the 16SCD files with the reference's 15 residues and bead 500.
 */
int tis_build_chain(const tis_chain_files *files, const tis_build_args *args, tis_system **out,
                    long long *rand_draws_out);

/* A seeded random-sequence chain of any length. */
typedef struct tis_synthetic_args {
    int n_nt;            /* nucleotides (>= 8); no 5' phosphate, as in 16SCD */
    double stack_frac;   /* probability that an adjacent nucleotide pair carries a secondary stack */
    int n_hb;            /* HB candidates wanted (fewer if the walk makes fewer base contacts) */
    double radius;       /* radius (A) of the sphere the walk grows in; box >= 5 * radius */
    tis_build_args build;/* ss_D hs_D st_D T c_mg c_k seed box, meaning as for tis_build_16scd */
} tis_synthetic_args;

/*
 * Builds a synthetic chain in place of the 16SCD input files: the reference's
 * set-up (read_PDB, initialize_HB_atoms, initialize_unprocessed_bonds/stacks,
 * fast.cpp:3137-4492) is hard-wired to 16SCD's residue numbering
 * (fast.cpp:3783-3795) and stack-face exceptions (fast.cpp:4458-4491); this
 * generalises it to n_nt nucleotides.  The chain is a self-avoiding walk grown
 * inside a sphere (bonds and bond angles at the backbone terms' equilibrium
 * values, no two other beads closer than 3.5 A), with secondary stacks on a
 * random stack_frac of adjacent pairs and HB candidates between bases that the
 * walk puts 5-7 A apart (each nucleotide in at most two).  Then, as for
 * 16SCD: velocities, set_box on the middle bead, ions by add_crowder.  Only
 * MAXIKEY_Mg_Cl_K.dat is read from data_dir.  All draws come from the glibc
 * stream seeded with srand((unsigned)(T*1000*seed)).
 */
int tis_build_synthetic(const char *data_dir, const tis_synthetic_args *args, tis_system **out);
void tis_system_free(tis_system *sys);

/* The glibc rand() stream (TYPE_3 additive generator) after srand(seed), values
 * skip .. skip+n-1.  The reference draws every random number from it; the host
 * uses it for set-up and to feed TIS_RNG_STREAM contexts. */
int tis_glibc_rand_fill(unsigned seed, long long skip, int32_t *out, long long n);

/* ------------------------------------------------------------ the engine */
typedef struct tis_ctx tis_ctx;

enum { TIS_PREC_FP64 = 0,   /* all force arithmetic in FP64 ("plain") */
       TIS_PREC_MIXED = 1 };/* pair / multi-body forces in FP32, state and sums in FP64 */

enum { TIS_RNG_PHILOX = 0,  /* counter-based Philox4x32-10 per particle and step */
       TIS_RNG_STREAM = 1 };/* caller-fed uniform stream in the reference's draw order */

typedef struct tis_options {
    int precision;        /* TIS_PREC_* */
    int rng;              /* TIS_RNG_* */
    uint64_t seed;        /* Philox key */
    double noise_scale;   /* multiplies SIGMA_FORCE; 1 = reference, 0 = noise-free */
    double skin;          /* list skin dr1 (fast.cpp:6227), 3 A in the reference */
    double r_sr;          /* short-range list radius, 12 + skin in the reference */
    double r_lr;          /* long-range list radius, 50 + skin in the reference */
    int lr_every;         /* long-range Coulomb refresh cadence, 10 in the reference */
    int use_graph;        /* capture each tis_step chunk in a hipGraph */
    int graph_steps;      /* steps per captured graph */
    long long exchange_every; /* steps between replica-exchange attempts; 0 = none */
    /* launch shapes (no counterpart in the reference; chosen by tis_autotune).
     * Results do not depend on them beyond the summation order of each
     * particle's pair forces. */
    int pair_lanes;       /* lanes per particle in the pair gather: power of two, 1..64 (4) */
    int pair_block;       /* workgroup size of the pair gather: multiple of 64, <= 256 (256) */
    int list_waves;       /* wavefronts of the neighbour-list build grid, 64..65536 (4096) */
    int hb_groups;        /* workgroups per replica sharing the components of the HB valence
                             resolution (Philox draws), 1..16 (4) */
    int hb_table_lag;     /* 1 (default): the HB tables of a list rebuild inside tis_step are
                             built beside that step's valence resolution, which reads the
                             previous rebuild's tables, and are used from the next step on.
                             Results are identical (every candidate within its gate is in
                             both); taken only when max r0 + 2 + skin + 1 < r_sr.  0: built
                             before the resolution. */
    int persistent;       /* 1: tis_step runs its steps as one cooperative launch per 20,000
                             steps in which each replica advances on its own workgroups (all
                             on one XCD), with barriers among them instead of kernel
                             boundaries, so replicas do not wait for each other; taken when
                             the context allows it (Philox draws, at least 8 replicas, see
                             tis_get_persistent_counts).  0 (default): the step's kernels
                             launched per step or replayed from a hipGraph, which is faster
                             on MI355X for the 16-replica batch (measured: 144 k against
                             112 k replica-steps/s). */
} tis_options;

/* Fills *opt with the reference's settings (FP64, stream RNG off => Philox, skin 3, ...). */
void tis_default_options(tis_options *opt);

/* Copies n_rep replica descriptions to the device. Replaces the set-up of
 * fast.cpp:6219-6239 (MPI_Init, list buffers, populate_lists). */
int tis_create(const tis_system *reps, int n_rep, const tis_options *opt, tis_ctx **out);
void tis_destroy(tis_ctx *ctx);

/* Autotuner (no counterpart in the reference).  Times tis_step on a scratch
 * context built from the same replicas and options (the caller's systems are
 * not modified) and picks, by coordinate descent, the fastest of
 *   graph_steps  25, 50, 100, 200    (steps per captured hipGraph)
 *   pair_lanes   2, 4, 8, 16         (lanes per particle in the pair gather)
 *   pair_block   64, 128, 256        (its workgroup size)
 *   list_waves   1024 .. 16384       (grid of the neighbour-list build)
 *   hb_groups    1, 2, 4, 8          (workgroups per replica in the HB resolution; Philox only)
 * writing them into *opt.  Each setting is timed twice over trial_steps steps
 * (after trial_steps / 4 warm-up steps) and its faster run is kept.  The list
 * skin, the list radii and the long-range cadence are not tuned: they decide
 * which pairs the reference treats as short- or long-range at each step, so
 * they are part of the model, not of its schedule. */
typedef struct tis_tuning {
    double ms_per_step_start;  /* the options passed in */
    double ms_per_step_best;   /* the options written back */
    int trials;                /* settings timed */
} tis_tuning;
int tis_autotune(const tis_system *reps, int n_rep, tis_options *opt, int trial_steps, tis_tuning *report);

/* State in/out (xyz-interleaved, 3*N doubles each; any pointer may be NULL on get).
 * set_state also rebuilds the pair lists (populate_lists, fast.cpp:1546-1737) and
 * sets PROGRESS = 0 and the step counter klok. */
int tis_set_state(tis_ctx *ctx, int rep, const double *pos, const double *pos_old,
                  const double *vel, long long klok);
int tis_get_state(tis_ctx *ctx, int rep, double *pos, double *pos_old, double *vel);
/* Sets the step counter klok only (lists, PROGRESS and the held long-range force
 * continue), as the second deck of the _DD driver does (DD.cpp:6428). */
int tis_set_step_counter(tis_ctx *ctx, int rep, long long klok);

/* Philox stream family of replica rep (rep = -1: every replica).  Draws are keyed
 * by (seed, replica, epoch) and counted by the step counter klok, so a driver
 * that restarts klok on the same context, as the second deck of the _DD driver
 * does (klok = 1 again, DD.cpp:6428), moves to another epoch and no draw of the
 * first deck repeats.  (The reference's single rand() stream just keeps
 * advancing across decks.)  0 at creation; tis_set_state leaves it unchanged. */
int tis_set_rng_epoch(tis_ctx *ctx, int rep, unsigned epoch);

/* Advances every replica n_steps Langevin steps: the body of fast.cpp:6274-6387
 * without its output block (deterministic_forces_MPI, full_forces,
 * move_rigid_units, check_shifts). */
int tis_step(tis_ctx *ctx, long long n_steps);

/* Captures and instantiates the hipGraph of graph_steps steps that tis_step
 * replays, without launching it, so that a later timed tis_step call holds no
 * capture; graph_steps replaces the context's option (0: launch every step
 * one by one).  No counterpart in the reference. */
int tis_prepare_graph(tis_ctx *ctx, int graph_steps);
/* The persistent step (tis_options.persistent; no counterpart in the
 * reference): on = 0 switches it off for later tis_step calls, on = 1 back on
 * (TIS_EINVAL when the context could not set it up).  Results follow the same
 * step order and differ from the separate kernels' only in the summation
 * order of atomically added forces. */
int tis_set_persistent(tis_ctx *ctx, int on);
/* Whether the persistent step is in use, its launches and steps since
 * creation, and (why, why_len) the reason when it is not available (empty
 * otherwise).  Any pointer may be NULL. */
int tis_get_persistent_counts(tis_ctx *ctx, int *active, long long *launches, long long *steps, char *why,
                              int why_len);
/* Since creation: graph replays, steps run inside them, steps launched one by
 * one (any pointer may be NULL). */
int tis_get_launch_counts(tis_ctx *ctx, long long *graph_launches, long long *graph_steps,
                          long long *eager_steps);

/* Phase trace of the HB valence resolution (profiling; no counterpart in the
 * reference, whose only timing code is commented out, fast.cpp:6520-6541).
 * on = 1: every later resolution writes, per (step slot klok % 64, replica,
 * group), 10 words: the device real-time clock (100 MHz ticks) at the group's
 * start and at the end of its phases (tables in LDS, excess ranked, components
 * joined, roots ranked, resolution, outcome and forces, end), the serial depth
 * (passes of its busiest wavefront), then n_over | n_comp << 16 | n_live << 32
 * (over-subscribed sites, components, live candidates).  Results are unchanged.
 * tis_get_hb_trace copies the ring, 64 * n_rep * 16 * 10 words (groups beyond
 * hb_groups stay 0), into out (n words available). */
int tis_trace_hb(tis_ctx *ctx, int on);
int tis_get_hb_trace(tis_ctx *ctx, unsigned long long *out, long long n);

/* Replica exchange (temperature ladder).  No counterpart in the reference,
 * whose replicas are separate MPI jobs; this is the on-device swap of the
 * batched configuration.  Neighbouring replicas r and r + 1 exchange when they
 * describe the same system (particle types, bonds, angles, stacks and HB
 * candidates equal; temperatures and set-up draws may differ).  An attempt, every
 * exchange_every steps of tis_step and on each tis_exchange_now, pairs
 * (r, r+1) with r even on even attempts and r odd on odd ones; for each pair it
 *   1. evaluates every replica's energy U_r(x_r) (all terms; the HB term after
 *      the valence resolution, which does not depend on T),
 *   2. swaps the pair's configurations (velocities and x - x_old scaled by
 *      sqrt(T_to / T_from)), rebuilds their lists and evaluates the non-HB
 *      energies U_r(x_s),
 *   3. accepts with probability min(1, exp(-D)),
 *      D = (U_r(x_s) - U_r(x_r)) / T_r + (U_s(x_r) - U_s(x_s)) / T_s
 *          + (H(x_s) - H(x_r)) (1/T_r - 1/T_s)
 *      (H = HB energy of each configuration from step 1), with a Philox draw;
 *   4. restores rejected pairs exactly.
 * Energies depend on T through lB(T) and the secondary-stack E(T). */
int tis_exchange_now(tis_ctx *ctx);
/* Per neighbour pair (r, r + 1), r < n_rep - 1: attempts, accepts and the last D
 * (0 where no attempt was made); any pointer may be NULL. */
int tis_get_exchange_stats(tis_ctx *ctx, long long *attempts, long long *accepts, double *last_d, int n_pairs);
/* Trajectory (walker) id held by each replica slot; initially walker[r] = r. */
int tis_get_walkers(tis_ctx *ctx, int *walker, int n_rep);

/* Force matching (no counterpart in the reference; the differentiable path of
 * the batched engine).  With F the forces a step uses (HB after the valence
 * resolution, whose outcome is held fixed for the derivatives; bonded and 1-3
 * pairs without WCA; the long-range term evaluated afresh) at every replica's
 * current state and target the caller's forces (3 * N_total doubles, replicas
 * concatenated),
 *   loss = (1/M) sum |F - target|^2,  M = 3 * N_total,
 * and, for each non-NULL output:
 *   forces     F (3 * N_total);
 *   grad_pos   dloss/dx (3 * N_total) = (2/M) J (F - target), J = dF/dx (a
 *              Hessian-vector product, by forward-mode derivatives of the force
 *              kernels);
 *   grad_hb_E  dloss/d(hb_E[k]) for the n_hb candidates (summed over replicas);
 *   grad_st3_E, grad_st2_E  dloss/d(E) of the tertiary / secondary stacks (the
 *              stacking free energy enters as E = -dG + ..., so dloss/d(dG) = -this);
 *   grad_ion_R dloss/d(R) for the LJ radii of Mg, Cl, K (through D_LJ = R_a + R_b).
 * Per-term gradients need replicas that describe one system (the snapshots of a
 * force-matching set).  Each replica's lists and HB tables are those of its
 * current state (tis_set_state rebuilds them). */
int tis_force_match(tis_ctx *ctx, const double *target, double *loss, double *forces, double *grad_pos,
                    double *grad_hb_E, double *grad_st3_E, double *grad_st2_E, double *grad_ion_R);
/* tis_force_match without any allocation.  No counterpart in the reference; for
 * callers that cannot allocate device memory at the call (serving loops, code
 * that forbids hipMalloc).  Every pointer is DEVICE memory the caller owns, laid
 * out as for tis_force_match (target, forces, grad_pos 3 * N_total doubles; loss
 * 1; per-term gradients n_hb, n_st3, n_st2 of one replica; grad_ion_R 3); target
 * and loss are required, the others may be NULL.  The working sums live in a
 * fixed array of the library's code object, 1 MiB per device (4 + n_hb + n_st3 +
 * n_st2 doubles at most 131,072, else TIS_EINVAL), held until the process ends
 * and shared by every context on that device: calls that overlap on one device
 * race.  The call returns after its work on the context's stream is done. */
int tis_force_match_noalloc(tis_ctx *ctx, const double *target, double *loss, double *forces, double *grad_pos,
                            double *grad_hb_E, double *grad_st3_E, double *grad_st2_E, double *grad_ion_R);

/* Replaces force-field parameters in every replica: HB strengths (n_hb), tertiary
 * and secondary stack energies (n_st3, n_st2) and the LJ radii of Mg, Cl, K (which
 * also set their friction, gamma = 0.555 R, as read_maxi_key does, fast.cpp:2897).
 * NULL leaves a set unchanged; per-term sets need replicas of one system. */
int tis_set_force_field(tis_ctx *ctx, const double *hb_E, const double *st3_E, const double *st2_E,
                        const double *ion_R);

/* Energies of the last force evaluation of replica rep. */
typedef struct tis_energies {
    double e2_hb, e3_hb, e2_stack, e3_stack;   /* as in the Energies file, fast.cpp:6312 */
    double e_wca, e_coul_sr, e_coul_lr;        /* pair terms (only after tis_eval_forces) */
    double e_bond, e_angle;                    /* backbone (only after tis_eval_forces) */
    double rg;                                 /* radius of gyration of the beads */
    long long klok;                            /* step counter */
    long long n_rebuilds;                      /* list rebuilds so far */
    double progress;                           /* PROGRESS */
    long long n_sr_pairs, n_lr_pairs, n_hb_active;
    long long hb_rounds;                       /* serial depth of the last HB valence resolution */
} tis_energies;
int tis_get_energies(tis_ctx *ctx, int rep, tis_energies *e);
/* Per-stack energies st_energy[1..s3_N] (the Contacts file, fast.cpp:6290). */
int tis_get_stack_energies(tis_ctx *ctx, int rep, double *out, int n);

/* Equilibrium observables, accumulated on the device.  The reference writes
 * frames (Coordinates / Energies / Contacts, fast.cpp:6276-6316) and computes
 * these offline from them; here every sample adds them in place:
 *   - per phosphate bead, the Mg, Cl and K ions within `cut` (minimum image;
 *     analysis/mg_correlation/mg_bound.F90:97-126) and whether at least one Mg
 *     is bound (count_mg_bound.py:8-15);
 *   - ions of each kind within `cut` of any phosphate (count_bound_total,
 *     analysis/salt_nt/dcd_traj_SaltCount_lB.py:52-62, where cut = r_kT = 7.2684 A);
 *   - chain shape from the gyration tensor of the beads: Rg, asphericity D and
 *     shape parameter S (analysis/shape/dcd_traj_shape.F90:193-226; Rg as
 *     radius_of_gyration, fast.cpp:2844-2876);
 *   - every energy term, and per HB candidate whether the valence resolution
 *     kept it, per tertiary stack its energy (the Contacts columns).
 * A sample evaluates every term at the current configuration (a pending list
 * rebuild is done first, as the next step would do; the trajectory does not
 * change). */
typedef struct tis_observables {
    long long n_samples;
    double cut;                                /* P-ion contact distance (A) */
    double rg, rg2, asphericity, shape;        /* means of Rg, Rg^2, D, S */
    double e2_hb, e3_hb, e2_stack, e3_stack;   /* means of the energy terms */
    double e_wca, e_coul_sr, e_coul_lr, e_bond, e_angle;
    double bound_mg, bound_cl, bound_k;        /* mean ions within cut of any P */
} tis_observables;
/* Clears the accumulators and sets the cadence: one sample every `every` steps
 * of tis_step (counted like exchange_every; 0 = samples only on tis_observe_now)
 * with P-ion contact distance `cut` (A; at most r_sr - skin - 1, so that every
 * pair within it is in the short-range lists). */
int tis_observe(tis_ctx *ctx, long long every, double cut);
/* One sample of every replica now. */
int tis_observe_now(tis_ctx *ctx);
/* Means over the samples of replica rep.  Optional arrays (NULL to skip):
 *   per_bead  4 * n_beads: for each phosphate bead the mean Mg, Cl, K counts
 *             within cut and the fraction of samples with >= 1 Mg (0 elsewhere);
 *             mean count / (4/3 pi cut^3 * 6.02214e-4) is the local
 *             concentration in mol/L (analysis/salt_nt/integrate_num_ions.py:51-55)
 *   hb_kept   n_hb: fraction of samples in which each candidate was kept
 *   st3_mean  n_st3: mean energy of each tertiary stack */
int tis_get_observables(tis_ctx *ctx, int rep, tis_observables *o, double *per_bead, int n_beads,
                        double *hb_kept, int n_hb, double *st3_mean, int n_st3);

/* Parity hook: forces of the selected terms at the current positions, without
 * moving anything.  HB uses the replica's uniform source (stream or Philox). */
enum { TIS_TERM_STACK = 1, TIS_TERM_HB = 2, TIS_TERM_BACKBONE = 4,
       TIS_TERM_SR = 8, TIS_TERM_LR = 16, TIS_TERM_ALL = 31 };
int tis_eval_forces(tis_ctx *ctx, int rep, unsigned term_mask, double *f_out, tis_energies *e);

/* Parity hook: current pair lists as (i,j) pairs with i<j, sorted; which 0 = SR, 1 = LR. */
int tis_get_pairs(tis_ctx *ctx, int rep, int which, int *ij_out, long long cap, long long *n_out);
/* Parity hook: hb_status (-1 / 0) of every candidate after the last list build. */
int tis_get_hb_status(tis_ctx *ctx, int rep, int *status_out, int n);
/* Parity hook: outcome of every candidate in the last HB evaluation: -1 not in the
 * SR list, 0 outside the distance gate, 1 kept, 2 switched off by the valence
 * resolution (hydrogen_bonds_interactions, fast.cpp:2165-2320).  "SR list" is the
 * latest list build; a candidate it adds after the last evaluation (HB tables
 * built beside a step's resolution, hb_table_lag) is outside its gate and reads 0. */
int tis_get_hb_outcome(tis_ctx *ctx, int rep, int *out, int n);

/* TIS_RNG_STREAM: raw glibc rand() values (0..RAND_MAX) consumed in the reference's
 * order (HB valence resolution, then 1500 Box-Muller pairs per step).  Resets the
 * replica's cursor to 0.  *consumed gives the number used so far. */
int tis_set_uniform_stream(tis_ctx *ctx, int rep, const int32_t *raw, long long n);
int tis_get_stream_cursor(tis_ctx *ctx, int rep, long long *consumed);

/* Kernel timing.  While profiling is on, tis_step runs without graph capture and
 * brackets every kernel launch with HIP events on the stream it is launched on;
 * each kernel runs alone (the device is idle before and after it), so the times
 * are those of the kernels by themselves, not of the overlapped step.
 * tis_kernel_times writes one line "name launches avg_ms" per kernel, accumulated
 * since profiling was switched on.  (No reference counterpart: the reference has
 * no live timing, only a commented-out ftime block at fast.cpp:6520-6541.) */
int tis_set_profiling(tis_ctx *ctx, int on);
int tis_kernel_times(tis_ctx *ctx, char *buf, int buflen);

/* Philox4x32-10 on the host, the generator the engine uses on the device
 * (known-answer tests). */
void tis_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif /* TIS_GPU_H */
