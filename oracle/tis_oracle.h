/*
 * tis_oracle.h - FP64 CPU restatement of the reference's Langevin step.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP engine: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (lib16s engine, host driver) never links or calls it.
 *
 * It restates, in plain C++ double arithmetic, the per-step routines of
 * src/16SCD_Mg_Cl_K_COPY_fast.cpp ("fast.cpp"):
 *   separate_in_cells / populate_lists   fast.cpp:1461-1737  -> oracle_set_state, rebuilds
 *   check_shifts                         fast.cpp:1741-1776
 *   SR_interactions_lists / LR_...       fast.cpp:1783-1907 (+ LJ_FORCE, COULOMB_TRUNCATED_FORCE)
 *   stacking_interactions                fast.cpp:1911-2141
 *   hydrogen_bonds_interactions          fast.cpp:2145-2350
 *   polymer_constraints                  fast.cpp:2354-2520
 *   deterministic_forces_MPI (1 rank)    fast.cpp:2524-2586
 *   full_forces                          fast.cpp:2590-2625
 *   move_rigid_units / leap_frog / adjust_box  fast.cpp:570-642, 2632-2732
 *   radius_of_gyration                   fast.cpp:2844-2876
 *   WCA_ENERGY / COULOMB_TRUNCATED_ENERGY fast.cpp:749-837 (energies the reference never sums)
 * It consumes the same tis_system description as the engine (include/tis_gpu.h)
 * but derives the backbone terms and secondary-stack switching from the unit and
 * face tables itself, as the reference does.
 */
#ifndef TIS_ORACLE_H
#define TIS_ORACLE_H

#include <stdint.h>

#include "../include/tis_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

/* energies[] layout of oracle_eval */
enum { OE_E2_HB, OE_E3_HB, OE_E2_STACK, OE_E3_STACK, OE_WCA, OE_COUL_SR, OE_COUL_LR,
       OE_BOND, OE_ANGLE, OE_RG, OE_N };

oracle_ctx *oracle_create(const tis_system *sys, double r_sr, double r_lr, double skin, int lr_every);
void oracle_destroy(oracle_ctx *o);
/* Sets x, x_old, v (3N interleaved), the step counter and rebuilds the lists
 * (cells unless brute != 0). */
int oracle_set_state(oracle_ctx *o, const double *pos, const double *pos_old, const double *vel,
                     long long klok, int brute);
int oracle_get_state(const oracle_ctx *o, double *pos, double *pos_old, double *vel);
/* Raw glibc rand() values consumed in the reference's order; resets the cursor. */
int oracle_set_stream(oracle_ctx *o, const int32_t *raw, long long n);
long long oracle_stream_cursor(const oracle_ctx *o);
/* Fills the stream with n values of the C library's rand() after srand(seed). */
int oracle_set_stream_glibc(oracle_ctx *o, unsigned seed, long long n);
/* Long runs: drops the consumed part of the stream and appends the next n
 * values of the same glibc sequence (the process's rand() state continues from
 * the last oracle_set_stream_glibc / oracle_stream_glibc_more; one context per
 * process at a time). */
int oracle_stream_glibc_more(oracle_ctx *o, long long n);
/* Forces of the selected terms (TIS_TERM_* mask) at the current state, written to
 * f_out (3N), energies to e_out[OE_N].  The HB term consumes the stream. */
int oracle_eval(oracle_ctx *o, unsigned mask, double *f_out, double *e_out);
/* Full Langevin steps; noise_scale multiplies SIGMA_FORCE (0 = noise-free but the
 * 3000 draws per step are still consumed, as in the reference). */
int oracle_step(oracle_ctx *o, long long n, double noise_scale);
/* Current lists as sorted (i<j) pairs; which 0 = SR, 1 = LR. */
long long oracle_pairs(const oracle_ctx *o, int which, int *out, long long cap);
/* hb_status after the last list build (-1 / 0). */
int oracle_hb_status(const oracle_ctx *o, int *out);
/* Diagnostics of the last steps. */
long long oracle_rebuilds(const oracle_ctx *o);
double oracle_progress(const oracle_ctx *o);
long long oracle_klok(const oracle_ctx *o);
/* Per-tertiary-stack energies of the last force evaluation (n_st3 values). */
int oracle_stack_energies(const oracle_ctx *o, double *out);
/* Outcome of every candidate in the last HB evaluation: -1 not in the SR list,
 * 0 outside the distance gate, 1 kept, 2 switched off by the valence resolution. */
int oracle_hb_outcome(const oracle_ctx *o, int *out);
/* Number of HB valence-resolution iterations in the last HB evaluation. */
long long oracle_hb_iterations(const oracle_ctx *o);

#ifdef __cplusplus
}
#endif
#endif
