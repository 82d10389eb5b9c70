// tis_run: drop-in replacement for the reference's main() (fast.cpp:6038-6516,
// and the two-deck driver of 16SCD_Mg_Cl_K_COPY_fast_DD.cpp:6257-6558 with -G).
//
//   ./tis_run -D0.850 -E3.975 -d6.5 -T0.615 -c0.005 -f0.050 -s1      (src/run_sample.sh)
//
// Same command line, same input files (read from the working directory, or from
// -I<dir>), same three output files with the same names and formats:
//   Contacts_16SCD_EPS_Mg_Cl_K_D..._s<seed>.dat     49 tertiary stack energies "%+.3f "
//   Coordinates_16SCD_EPS_Mg_Cl_K_D..._s<seed>.dat  raw double x[0..N], y[0..N], z[0..N]
//   Energies_16SCD_EPS_Mg_Cl_K_D..._s<seed>.dat     "%le %le %le %le %le": E2_HB E3_HB
//                                                   E2_STACK E3_STACK Rg
// written every 10^6 steps before that step's forces, rolling copies
// "..._<klok/5e6>.dat" every 5*10^6 steps with the copy three intervals older
// removed, and restart from the last whole frame of an existing Coordinates file.
//
// Extra flags (absent = the reference's hard-coded values):
//   -L<box A>  -N<steps, 3e9>  -W<output interval, 1e6>  -P<64|32 force precision>
//   -R<philox|glibc>  -G (run the second deck of the _DD variant)  -I<input dir>
//   -S<Philox seed>  -B<steps per device call, 10000>
//   -O<sample interval>  equilibrium observables accumulated on the device every
//            that many steps and written at each output step as means over the
//            window since the previous one (the reference's offline analysis,
//            analysis/mg_correlation and analysis/shape, without frames):
//     Observables_16SCD_EPS_Mg_Cl_K_D..._s<seed>.dat  "klok samples Rg Rg^2 D S
//            bound_Mg bound_Cl bound_K E2_HB E3_HB E2_STACK E3_STACK"
//     MgBound_16SCD_EPS_Mg_Cl_K_D..._s<seed>.dat      per phosphate, the fraction of
//            samples with a Mg ion within r_kT = 7.2684 A ("%.4f ", chain order)
// The slot x[0] of each Coordinates frame is written as 0 (the reference leaves it
// uninitialised); partial trailing frames are rejected on restart.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "../../include/tis_gpu.h"
#include "glibc_rand.h"

namespace {

struct Args {
    double ss_D = 0, hs_D = 1.0, st_D = 0, T = 0, c_mg = 0, c_k = 0;
    int seed = 0;
    double box = 350.0;
    long long nsteps = 3000000000LL;
    long long write_every = 1000000;
    int precision = TIS_PREC_FP64;
    bool glibc = false;
    bool deck2 = false;
    std::string input_dir = ".";
    unsigned long long philox_seed = 0;
    long long chunk = 10000;
    long long observe_every = 0;
};

std::string tag(const Args &a, const char *seed_prefix) {
    char b[256];
    std::snprintf(b, sizeof b, "D%.3f_E%.3f_d%.3f_T%.3f_cM%.4f_f%.3f_s%s%d", a.ss_D, a.hs_D, a.st_D, a.T, a.c_mg,
                  a.c_k, seed_prefix, a.seed);
    return b;
}

constexpr double kRkT = 7.2684;   // P-ion contact distance of the reference's analysis

int die(const char *what, int rc) {
    std::fprintf(stderr, "tis_run: %s failed (%d): %s\n", what, rc, tis_last_error());
    return 1;
}

// Feeds a TIS_RNG_STREAM context with the glibc stream, continuing after the
// set-up draws, in chunks large enough for the next device call.
struct StreamFeeder {
    tis_host::GlibcRand g;
    std::vector<int32_t> pending;
    explicit StreamFeeder(unsigned seed, long long skip) : g(seed) {
        for (long long q = 0; q < skip; ++q) g.next();
    }
    int top_up(tis_ctx *ctx, long long need) {
        while ((long long)pending.size() < need) pending.push_back(g.next());
        return tis_set_uniform_stream(ctx, 0, pending.data(), (long long)pending.size());
    }
    int consume(tis_ctx *ctx) {
        long long used = 0;
        int rc = tis_get_stream_cursor(ctx, 0, &used);
        if (rc) return rc;
        pending.erase(pending.begin(), pending.begin() + used);
        return 0;
    }
};

struct Output {
    FILE *contacts = nullptr, *coords = nullptr, *energies = nullptr, *obs = nullptr, *mgb = nullptr;
    std::string n_contacts, n_coords, n_energies;
    ~Output() {
        for (FILE *f : {contacts, coords, energies, obs, mgb})
            if (f) std::fclose(f);
    }
};

void rolling_copy(const std::string &src, const std::string &stem, long long idx, bool remove_old) {
    namespace fs = std::filesystem;
    std::error_code ec;
    fs::copy_file(src, stem + "_" + std::to_string(idx) + ".dat", fs::copy_options::overwrite_existing, ec);
    if (remove_old) fs::remove(stem + "_" + std::to_string(idx - 3) + ".dat", ec);
}

}  // namespace

int main(int argc, char **argv) {
    Args a;
    for (int i = 1; i < argc; ++i) {
        if (argv[i][0] != '-') continue;
        const char *v = argv[i] + 2;
        switch (argv[i][1]) {
            case 'D': a.ss_D = std::atof(v); break;
            case 'E': a.hs_D = std::atof(v); break;
            case 'd': a.st_D = std::atof(v); break;
            case 'T': a.T = std::atof(v); break;
            case 'c': a.c_mg = std::atof(v); break;
            case 'f': a.c_k = std::atof(v); break;
            case 's': a.seed = std::atoi(v); break;
            case 'L': a.box = std::atof(v); break;
            case 'N': a.nsteps = std::atoll(v); break;
            case 'W': a.write_every = std::atoll(v); break;
            case 'P': a.precision = std::atoi(v) == 32 ? TIS_PREC_MIXED : TIS_PREC_FP64; break;
            case 'R': a.glibc = std::strcmp(v, "glibc") == 0; break;
            case 'G': a.deck2 = true; break;
            case 'I': a.input_dir = v; break;
            case 'S': a.philox_seed = std::strtoull(v, nullptr, 10); break;
            case 'B': a.chunk = std::max(1LL, std::atoll(v)); break;
            case 'O': a.observe_every = std::max(0LL, std::atoll(v)); break;
            default: std::printf("unrecognized argument %s\n", argv[i]); break;
        }
    }
    if (a.write_every <= 0) a.write_every = 1000000;
    tis_build_args ba{a.ss_D, a.hs_D, a.st_D, a.T, a.c_mg, a.c_k, a.seed, a.box};
    tis_system *sys = nullptr;
    long long draws = 0;
    int rc = tis_build_16scd(a.input_dir.c_str(), &ba, &sys, &draws);
    if (rc) return die("tis_build_16scd", rc);
    const int N = sys->n_particles;
    const std::string t1 = tag(a, "");
    Output out;
    out.n_contacts = "Contacts_16SCD_EPS_Mg_Cl_K_" + t1 + ".dat";
    out.n_coords = "Coordinates_16SCD_EPS_Mg_Cl_K_" + t1 + ".dat";
    out.n_energies = "Energies_16SCD_EPS_Mg_Cl_K_" + t1 + ".dat";
    // restart: last whole frame of an existing Coordinates file (fast.cpp:6181-6209)
    long long klok_old = 0;
    const size_t frame = 3 * (size_t)(N + 1) * sizeof(double);
    if (FILE *f = std::fopen(out.n_coords.c_str(), "rb")) {
        std::fseek(f, 0, SEEK_END);
        const long long bytes = std::ftell(f);
        if (bytes % (long long)frame != 0) {
            std::fprintf(stderr, "tis_run: %s holds a partial frame (%lld bytes, frame %zu); not restarting\n",
                         out.n_coords.c_str(), bytes, frame);
            std::fclose(f);
            tis_system_free(sys);
            return 1;
        }
        const long long frames = bytes / (long long)frame;
        if (frames > 0) {
            std::vector<double> buf(3 * (size_t)(N + 1));
            std::fseek(f, (long)((frames - 1) * (long long)frame), SEEK_SET);
            if (std::fread(buf.data(), 1, frame, f) != frame) {
                std::fclose(f);
                tis_system_free(sys);
                return die("reading the restart frame", TIS_EIO);
            }
            for (int i = 0; i < N; ++i)
                for (int c = 0; c < 3; ++c) {
                    sys->pos[3 * i + c] = buf[(size_t)c * (N + 1) + 1 + i];
                    sys->pos_old[3 * i + c] = sys->pos[3 * i + c] - sys->vel[3 * i + c] * sys->h;
                }
            klok_old = frames * a.write_every;
        }
        std::fclose(f);
    }
    tis_options opt;
    tis_default_options(&opt);
    opt.precision = a.precision;
    opt.rng = a.glibc ? TIS_RNG_STREAM : TIS_RNG_PHILOX;
    opt.seed = a.philox_seed ? a.philox_seed : (unsigned long long)(a.T * 1000 * a.seed);
    tis_ctx *ctx = nullptr;
    rc = tis_create(sys, 1, &opt, &ctx);
    if (rc) { tis_system_free(sys); return die("tis_create", rc); }
    std::vector<double> pos(3 * (size_t)N), st(sys->n_st3 > 0 ? sys->n_st3 : 1);
    std::vector<double> frame_buf(3 * (size_t)(N + 1));
    StreamFeeder feeder((unsigned)(a.T * 1000 * a.seed), draws);
    const int n_st3 = sys->n_st3;
    const int n_beads = sys->n_beads;
    std::vector<double> per_bead(4 * (size_t)n_beads);
    std::vector<int> p_beads;   // phosphate beads in chain order
    for (int i = 0; i < n_beads; ++i)
        if (sys->cls[i] == 0) p_beads.push_back(i);

    auto run_deck = [&](const char *seed_prefix, long long klok_first) -> int {
        const std::string t = tag(a, seed_prefix);
        Output o;
        o.n_contacts = "Contacts_16SCD_EPS_Mg_Cl_K_" + t + ".dat";
        o.n_coords = "Coordinates_16SCD_EPS_Mg_Cl_K_" + t + ".dat";
        o.n_energies = "Energies_16SCD_EPS_Mg_Cl_K_" + t + ".dat";
        o.contacts = std::fopen(o.n_contacts.c_str(), "a");
        o.coords = std::fopen(o.n_coords.c_str(), "ab");
        o.energies = std::fopen(o.n_energies.c_str(), "a");
        if (!o.contacts || !o.coords || !o.energies) {
            std::fprintf(stderr, "tis_run: cannot open the output files\n");
            return 1;
        }
        int r = tis_set_step_counter(ctx, 0, klok_first);
        if (r) return die("tis_set_step_counter", r);
        if (a.observe_every > 0) {
            o.obs = std::fopen(("Observables_16SCD_EPS_Mg_Cl_K_" + t + ".dat").c_str(), "a");
            o.mgb = std::fopen(("MgBound_16SCD_EPS_Mg_Cl_K_" + t + ".dat").c_str(), "a");
            if (!o.obs || !o.mgb) {
                std::fprintf(stderr, "tis_run: cannot open the observables files\n");
                return 1;
            }
            if ((r = tis_observe(ctx, a.observe_every, kRkT))) return die("tis_observe", r);
        }
        long long klok = klok_first;
        while (klok <= a.nsteps) {
            if (klok % a.write_every == 0) {
                tis_energies e;
                if ((r = tis_get_energies(ctx, 0, &e))) return die("tis_get_energies", r);
                if ((r = tis_get_stack_energies(ctx, 0, st.data(), n_st3))) return die("tis_get_stack_energies", r);
                if ((r = tis_get_state(ctx, 0, pos.data(), nullptr, nullptr))) return die("tis_get_state", r);
                double e3 = 0;
                for (int k = 0; k < n_st3; ++k) {
                    std::fprintf(o.contacts, "%+.3f ", st[k]);
                    e3 += st[k];
                }
                std::fprintf(o.contacts, "\n");
                std::fflush(o.contacts);
                for (int c = 0; c < 3; ++c) {
                    frame_buf[(size_t)c * (N + 1)] = 0.0;
                    for (int i = 0; i < N; ++i) frame_buf[(size_t)c * (N + 1) + 1 + i] = pos[3 * i + c];
                }
                std::fwrite(frame_buf.data(), sizeof(double), frame_buf.size(), o.coords);
                std::fflush(o.coords);
                std::fprintf(o.energies, "%le %le %le %le %le\n", e.e2_hb, e.e3_hb, e.e2_stack, e3, e.rg);
                std::fflush(o.energies);
                if (o.obs) {
                    tis_observables ob;
                    if ((r = tis_get_observables(ctx, 0, &ob, per_bead.data(), n_beads, nullptr, 0, nullptr, 0)))
                        return die("tis_get_observables", r);
                    if (ob.n_samples > 0) {
                        std::fprintf(o.obs, "%lld %lld %le %le %le %le %le %le %le %le %le %le %le\n", klok,
                                     ob.n_samples, ob.rg, ob.rg2, ob.asphericity, ob.shape, ob.bound_mg, ob.bound_cl,
                                     ob.bound_k, ob.e2_hb, ob.e3_hb, ob.e2_stack, ob.e3_stack);
                        for (int b : p_beads) std::fprintf(o.mgb, "%.4f ", per_bead[4 * (size_t)b + 3]);
                        std::fprintf(o.mgb, "\n");
                        std::fflush(o.obs);
                        std::fflush(o.mgb);
                    }
                    if ((r = tis_observe(ctx, a.observe_every, kRkT))) return die("tis_observe", r);   // next window
                }
                if (klok % (5 * a.write_every) == 0 || klok == a.nsteps) {
                    const long long idx = klok / (5 * a.write_every);
                    const bool last = klok == a.nsteps;
                    const std::string sc = "Contacts_16SCD_EPS_Mg_Cl_K_" + t;
                    const std::string sx = "Coordinates_16SCD_EPS_Mg_Cl_K_" + t;
                    const std::string se = "Energies_16SCD_EPS_Mg_Cl_K_" + t;
                    rolling_copy(o.n_contacts, sc, idx, !last);
                    rolling_copy(o.n_coords, sx, idx, !last);
                    rolling_copy(o.n_energies, se, idx, !last);
                }
            }
            // steps up to (not including) the next output step, at most one chunk
            const long long next_out = (klok / a.write_every + 1) * a.write_every;
            const long long n = std::min({next_out - klok, a.nsteps + 1 - klok, a.chunk});
            if (a.glibc && (r = feeder.top_up(ctx, n * 6000 + 10000))) return die("tis_set_uniform_stream", r);
            if ((r = tis_step(ctx, n))) return die("tis_step", r);
            if (a.glibc && (r = feeder.consume(ctx))) return die("tis_get_stream_cursor", r);
            klok += n;
        }
        return 0;
    };
    rc = run_deck("", klok_old + 1);
    // deck 2 (DD.cpp:6405-6544) continues from the final state with klok = 1; its
    // Philox draws come from epoch 1, so the restarted step counter repeats none of
    // deck 1's (the glibc stream just continues, as the reference's rand() does)
    if (!rc && a.deck2) {
        if ((rc = tis_set_rng_epoch(ctx, 0, 1))) rc = die("tis_set_rng_epoch", rc);
        else rc = run_deck("1", 1);
    }
    tis_destroy(ctx);
    tis_system_free(sys);
    return rc;
}
