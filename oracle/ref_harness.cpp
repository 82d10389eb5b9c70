// Test-only harness around the unmodified reference translation unit.
//
// This file is part of the oracle (test infrastructure).  It is compiled by
// oracle/Makefile *only* where the reference sources are present, and it never
// ships in the product.  The reference source is included from where it lies
// (REF_SRC is given on the compiler command line); nothing is copied.
//
// What it does: runs the reference's own initialisation routines in the order
// of its main() (src/16SCD_Mg_Cl_K_COPY_fast.cpp:6105-6239), with a settable box
// and enlarged pair-list buffers, then dumps the resulting topology tables,
// initial state, pair lists, per-term forces and a short trajectory into one
// binary record file.  The glibc rand() calls made by the reference are counted
// through a macro wrapper so that the oracle can replay the same stream.
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <mpi.h>
#include <ctime>
#include <vector>

static long long g_rand_calls = 0;
static int harness_rand(void) { ++g_rand_calls; return rand(); }

#define rand harness_rand
#define main tis_reference_main
#include REF_SRC
#undef main
#undef rand

namespace {

FILE *g_out = nullptr;

void rec(const char *name, char kind, long long n, const void *p) {
    char tag[32];
    memset(tag, 0, sizeof tag);
    strncpy(tag, name, 31);
    fwrite(tag, 1, 32, g_out);
    fwrite(&kind, 1, 1, g_out);
    fwrite(&n, sizeof n, 1, g_out);
    size_t w = (kind == 'd' || kind == 'l') ? 8 : 4;
    if (n) fwrite(p, w, (size_t)n, g_out);
}
void rec_i(const char *name, long long n, const int *p) { rec(name, 'i', n, p); }
void rec_d(const char *name, long long n, const double *p) { rec(name, 'd', n, p); }
void rec_scalar_d(const char *name, double v) { rec_d(name, 1, &v); }
void rec_scalar_l(const char *name, long long v) { rec(name, 'l', 1, &v); }

void dump_xyz(const char *pfx, const double *a, const double *b, const double *c) {
    char nm[32];
    snprintf(nm, sizeof nm, "%s_x", pfx); rec_d(nm, NT_atom, a + 1);
    snprintf(nm, sizeof nm, "%s_y", pfx); rec_d(nm, NT_atom, b + 1);
    snprintf(nm, sizeof nm, "%s_z", pfx); rec_d(nm, NT_atom, c + 1);
}

void zero_f(double *a, double *b, double *c) {
    for (int i = 0; i <= NT_atom; ++i) { a[i] = 0; b[i] = 0; c[i] = 0; }
}

// CSR flattening of the reference's ragged 1-based tables.
void dump_csr(const char *name, int n_rows, const int *row_n, int *const *rows) {
    std::vector<int> cnt(n_rows), flat;
    for (int r = 1; r <= n_rows; ++r) {
        cnt[r - 1] = row_n[r];
        for (int q = 1; q <= row_n[r]; ++q) flat.push_back(rows[r][q]);
    }
    char nm[32];
    snprintf(nm, sizeof nm, "%s_n", name); rec_i(nm, n_rows, cnt.data());
    snprintf(nm, sizeof nm, "%s_v", name); rec_i(nm, (long long)flat.size(), flat.data());
}

}  // namespace

int main(int argc, char **argv) {
    double box = 350.0;
    long nsteps = 0, every = 1;
    int noise_off = 0, timing = 0;
    unsigned rng_terms = 12345u, rng_traj = 777u;
    const char *out = "ref_dump.bin";
    int seed = 1;
    for (int a = 1; a < argc; ++a) {
        if (argv[a][0] != '-') continue;
        const char *v = argv[a] + 2;
        switch (argv[a][1]) {
            case 'D': ss_D = atof(v); break;
            case 'E': hs_D = atof(v); break;
            case 'd': st_D = atof(v); break;
            case 'T': T = atof(v); break;
            case 'c': cM_crwd[0] = atof(v); break;
            case 'f': cM_crwd[2] = atof(v); break;
            case 's': seed = atoi(v); break;
            case 'L': box = atof(v); break;
            case 'N': nsteps = atol(v); break;
            case 'K': every = atol(v); break;
            case 'Z': noise_off = 1; break;
            case 'R': rng_terms = (unsigned)atol(v); break;
            case 'Q': rng_traj = (unsigned)atol(v); break;
            case 'O': out = v; break;
            case 'P': timing = 1; break;
            default: fprintf(stderr, "unknown flag %s\n", argv[a]); return 2;
        }
    }
    side_X = box; side_Y = box; side_Z = box;
    srand((unsigned)(T * 1000 * seed));
    lB = 332.0637090 * (0.03273600947 * T - 0.00669625750);
    read_maxi_key((char *)"MAXIKEY_Mg_Cl_K.dat");
    list_crowder_types();
    read_PDB((char *)"16SCD.pdb");
    initialize_HB_atoms();
    initialize_unprocessed_bonds((char *)"16SCD_16SCD_unprocessed_bonds.dat");
    initialize_unprocessed_stacks((char *)"16SCD_16SCD_unprocessed_stacks.dat");
    set_box(500);
    N_crwd[0] = (int)(cM_crwd[0] * 6.022 * 0.0001 * side_XYZ);
    N_crwd[2] = (int)(cM_crwd[2] * 6.022 * 0.0001 * side_XYZ);
    N_crwd[1] = 2 * N_crwd[0] + N_crwd[2] - (NBP_atom - 1) / 2;
    for (int t = 0; t < n_crwd; ++t)
        for (int q = 1; q <= N_crwd[t]; ++q) add_crowder(t);
    long long rand_after_init = g_rand_calls;

    fsend = (double *)calloc(3 * NT_atom, sizeof(double));
    freceive = (double *)calloc(3 * NT_atom, sizeof(double));
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &my_rank);
    MPI_Comm_size(MPI_COMM_WORLD, &comm_size);
    dr1 = 3.0;
    R1_LIST[0] = 12.0 + dr1; R2_LIST[0] = R1_LIST[0] * R1_LIST[0];
    R1_LIST[1] = D_CT_CUTOFF + dr1; R2_LIST[1] = R1_LIST[1] * R1_LIST[1];
    // Enlarged buffers: the reference's 100001/500001 overflow above ~4k particles.
    size_t cap = (size_t)NT_atom * (size_t)NT_atom + 16;
    if (cap > (size_t)60000000) cap = 60000000;
    list_content[0] = (int *)calloc(cap, sizeof(int));
    list_content[1] = (int *)calloc(cap, sizeof(int));
    populate_lists();
    PROGRESS = 0;

    // -P: time nsteps of the reference's own step sequence (fast.cpp:6380-6386)
    // and print "steps <n> seconds <s>"; nothing is dumped
    if (timing) {
        srand(rng_traj);
        timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        for (klok = 1; klok <= nsteps; ++klok) {
            deterministic_forces_MPI();
            full_forces();
            move_rigid_units();
            check_shifts();
        }
        clock_gettime(CLOCK_MONOTONIC, &t1);
        printf("steps %ld seconds %.9f particles %d\n", nsteps,
               (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec), NT_atom);
        MPI_Finalize();
        return 0;
    }

    g_out = fopen(out, "wb");
    if (!g_out) { perror(out); return 1; }
    // ---- scalars
    rec_scalar_l("NT_atom", NT_atom);
    rec_scalar_l("NTP_atom", NTP_atom);
    rec_scalar_l("NBP_atom", NBP_atom);
    rec_scalar_l("hb_N", hb_N);
    rec_scalar_l("HB_NT_atom", HB_NT_atom);
    rec_scalar_l("s3_N", s3_N);
    rec_scalar_l("st_N", st_N);
    rec_scalar_l("N_Mg", N_crwd[0]);
    rec_scalar_l("N_Cl", N_crwd[1]);
    rec_scalar_l("N_K", N_crwd[2]);
    rec_scalar_l("rand_after_init", rand_after_init);
    rec_scalar_d("side_X", side_X);
    rec_scalar_d("T", T);
    rec_scalar_d("lB", lB);
    rec_scalar_d("h", h);
    // ---- per-type tables
    rec_d("MASS", na, MASS + 1);
    rec_d("RLJ", na, RLJ + 1);
    rec_d("ELJ", na, ELJ + 1);
    rec_d("QC", na, QC + 1);
    rec_d("K1", na, K1 + 1);
    rec_d("K2", na, K2 + 1);
    rec_d("SIGMA_FORCE", na, SIGMA_FORCE + 1);
    rec_d("D_LJ", npa, D_LJ + 1);
    rec_d("E_LJ", npa, E_LJ + 1);
    rec_d("F_LJ", npa, F_LJ + 1);
    rec_d("Q_C", npa, Q_C + 1);
    rec_d("D2_LJ_OVERLAP", npa, D2_LJ_OVERLAP + 1);
    rec_d("D2_LJ_MINIMUM", npa, D2_LJ_MINIMUM + 1);
    // ---- particles
    rec_i("maxi_key", NT_atom, maxi_key + 1);
    rec_i("part_key", NT_atom, part_key + 1);
    dump_xyz("pos", x, y, z);
    dump_xyz("old", x_old, y_old, z_old);
    dump_xyz("vel", vx, vy, vz);
    // ---- units (nucleosides odd, phosphates even)
    {
        std::vector<int> a0(NBP_atom), a1(NBP_atom), ak(NBP_atom);
        for (int u = 1; u <= NBP_atom; ++u) {
            a0[u - 1] = atom_key[u][0];
            a1[u - 1] = NS_atom[u] >= 1 ? atom_key[u][1] : 0;
            ak[u - 1] = amino_key[u];
        }
        rec_i("atom_key0", NBP_atom, a0.data());
        rec_i("atom_key1", NBP_atom, a1.data());
        rec_i("amino_key", NBP_atom, ak.data());
    }
    // ---- hydrogen bonds
    rec_i("hb_k10", hb_N, hb_k10 + 1); rec_i("hb_k11", hb_N, hb_k11 + 1);
    rec_i("hb_k12", hb_N, hb_k12 + 1); rec_i("hb_k20", hb_N, hb_k20 + 1);
    rec_i("hb_k21", hb_N, hb_k21 + 1); rec_i("hb_k22", hb_N, hb_k22 + 1);
    rec_d("hb_r", hb_N, hb_r + 1); rec_d("hb_theta1", hb_N, hb_theta1 + 1);
    rec_d("hb_theta2", hb_N, hb_theta2 + 1); rec_d("hb_psi", hb_N, hb_psi + 1);
    rec_d("hb_psi1", hb_N, hb_psi1 + 1); rec_d("hb_psi2", hb_N, hb_psi2 + 1);
    rec_d("hb_E", hb_N, hb_E + 1);
    rec_i("hb_dode", hb_N, hb_dode + 1);
    rec_i("VALENCE", HB_NT_atom, VALENCE + 1);
    dump_csr("ATOM_HB", hb_N, ATOM_HB_N, ATOM_HB);
    dump_csr("HB_ATOM", HB_NT_atom, HB_ATOM_N, HB_ATOM);
    // ---- stacks
    rec_i("st_i", st_N, st_i + 1);
    rec_i("st_j", s3_N, st_j + 1);
    rec_d("st_r", st_N, st_r + 1);
    rec_d("st_theta1", s3_N, st_theta1 + 1); rec_d("st_theta2", s3_N, st_theta2 + 1);
    rec_d("st_psi", s3_N, st_psi + 1); rec_d("st_psi1", s3_N, st_psi1 + 1);
    rec_d("st_psi2", s3_N, st_psi2 + 1);
    rec_d("st_E", st_N, st_E + 1);
    {
        std::vector<int> f0(st_N), f1(st_N);
        for (int k = 1; k <= st_N; ++k) { f0[k - 1] = ATOM_ST[k][0]; f1[k - 1] = ATOM_ST[k][1]; }
        rec_i("ATOM_ST0", st_N, f0.data());
        rec_i("ATOM_ST1", st_N, f1.data());
    }
    rec_i("ST_EXCESS", NBP_atom, ST_EXCESS + 1);
    // ---- lists at the initial state
    rec_i("list_SR", list_mass[0], list_content[0] + 1);
    rec_i("list_LR", list_mass[1], list_content[1] + 1);
    rec_i("hb_status0", hb_N, hb_status + 1);
    rec_i("HB_EXCESS0", HB_NT_atom, HB_EXCESS + 1);

    // ---- per-term forces at the initial state
    zero_f(fx, fy, fz);
    E2_STACK = E3_STACK = 0;
    stacking_interactions();
    dump_xyz("F_stack", fx, fy, fz);
    rec_scalar_d("E2_STACK", E2_STACK); rec_scalar_d("E3_STACK", E3_STACK);
    rec_d("st_energy", st_N, st_energy + 1);
    rec_i("st_status", st_N, st_status + 1);

    zero_f(fx, fy, fz);
    E2_HB = E3_HB = 0;
    srand(rng_terms);
    g_rand_calls = 0;
    hydrogen_bonds_interactions();
    dump_xyz("F_hb", fx, fy, fz);
    rec_scalar_l("hb_rng_seed", rng_terms);
    rec_scalar_l("hb_rand_calls", g_rand_calls);
    rec_scalar_d("E2_HB", E2_HB); rec_scalar_d("E3_HB", E3_HB);
    rec_d("hb_energy", hb_N, hb_energy + 1);
    rec_i("hb_status1", hb_N, hb_status + 1);

    zero_f(fx, fy, fz);
    polymer_constraints();
    dump_xyz("F_poly", fx, fy, fz);

    zero_f(fx, fy, fz);
    SR_interactions_lists();
    dump_xyz("F_SR", fx, fy, fz);

    zero_f(flx, fly, flz);
    LR_interactions_lists();
    dump_xyz("F_LR", flx, fly, flz);

    // Pair energies with the reference's own (otherwise unused) energy routines.
    {
        double e_wca = 0, e_coul_sr = 0, e_coul_lr = 0;
        for (int l = 0; l < 2; ++l) {
            for (int k = 1; k < list_mass[l] + 1; k += 2) {
                int i = list_content[l][k], j = list_content[l][k + 1];
                int a = maxi_key[i], b = maxi_key[j];
                int jn = a < b ? (a - 1) * na - (a - 1) * a / 2 + b : (b - 1) * na - (b - 1) * b / 2 + a;
                double xx[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
                half_shift(xx);
                double r2 = xx[0] * xx[0] + xx[1] * xx[1] + xx[2] * xx[2];
                bool charged = part_key[i] != 1 && part_key[i] != 2 && part_key[j] != 1 && part_key[j] != 2;
                if (l == 0) {
                    e_wca += WCA_ENERGY(r2, jn);
                    if (charged) e_coul_sr += COULOMB_TRUNCATED_ENERGY(r2, jn);
                } else {
                    e_coul_lr += COULOMB_TRUNCATED_ENERGY(r2, jn);
                }
            }
        }
        rec_scalar_d("E_WCA_SR", e_wca);
        rec_scalar_d("E_COUL_SR", e_coul_sr);
        rec_scalar_d("E_COUL_LR", e_coul_lr);
    }
    rec_scalar_d("Rg0", radius_of_gyration(1, NTP_atom));

    // ---- trajectory: the reference's step order (fast.cpp:6380-6386)
    if (nsteps > 0) {
        if (noise_off)
            for (int t = 1; t <= na; ++t) SIGMA_FORCE[t] = 0.0;
        srand(rng_traj);
        g_rand_calls = 0;
        rec_scalar_l("traj_rng_seed", rng_traj);
        rec_scalar_l("traj_noise_off", noise_off);
        rec_scalar_l("traj_nsteps", nsteps);
        rec_scalar_l("traj_every", every);
        std::vector<long long> draws(nsteps);
        std::vector<int> rebuilt(nsteps);
        std::vector<double> prog(nsteps), e2hb(nsteps), e3hb(nsteps), e2st(nsteps), e3st(nsteps);
        for (klok = 1; klok <= nsteps; ++klok) {
            deterministic_forces_MPI();
            full_forces();
            move_rigid_units();
            check_shifts();
            long s = klok - 1;
            draws[s] = g_rand_calls;
            rebuilt[s] = PROGRESS == 0 ? 1 : 0;
            prog[s] = PROGRESS;
            e2hb[s] = E2_HB; e3hb[s] = E3_HB; e2st[s] = E2_STACK; e3st[s] = E3_STACK;
            if (klok % every == 0) {
                char pfx[32];
                snprintf(pfx, sizeof pfx, "t%ld_pos", (long)klok); dump_xyz(pfx, x, y, z);
                snprintf(pfx, sizeof pfx, "t%ld_old", (long)klok); dump_xyz(pfx, x_old, y_old, z_old);
                snprintf(pfx, sizeof pfx, "t%ld_vel", (long)klok); dump_xyz(pfx, vx, vy, vz);
            }
        }
        rec("traj_draws", 'l', nsteps, draws.data());
        rec_i("traj_rebuilt", nsteps, rebuilt.data());
        rec_d("traj_progress", nsteps, prog.data());
        rec_d("traj_E2_HB", nsteps, e2hb.data());
        rec_d("traj_E3_HB", nsteps, e3hb.data());
        rec_d("traj_E2_STACK", nsteps, e2st.data());
        rec_d("traj_E3_STACK", nsteps, e3st.data());
    }
    fclose(g_out);
    MPI_Finalize();
    return 0;
}
