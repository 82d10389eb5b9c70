// Host side of the MI355X engine: device buffers, launches and the C ABI of
// include/tis_gpu.h.  One context drives a batch of replicas on one GPU; all
// work goes to the context's streams.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/tis_gpu.h"
#include "backward.h"
#include "observe.h"
#include "persist.h"

namespace tis_host {
void set_error(const std::string &msg);
}

using namespace tis;

namespace {

struct DeviceError {
    std::string msg;
};

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw DeviceError{std::string(#expr) + ": " + hipGetErrorString(e_)};               \
    } while (0)

struct Fail {
    int code;
    std::string msg;
};

template <class T>
T *dalloc(std::vector<void *> &owned, size_t n) {
    void *p = nullptr;
    HIP_TRY(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    HIP_TRY(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T)));
    owned.push_back(p);
    return (T *)p;
}

template <class T>
void upload(T *dst, const std::vector<T> &src) {
    if (!src.empty()) HIP_TRY(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
}

}  // namespace

// Working memory of tis_force_match_noalloc, in the code object so that the
// entry point allocates nothing: its loss and ion-radius sums and the per-term
// gradient sums (4 + n_hb + n_st3 + n_st2 doubles, at most kScratchDoubles,
// 1 MiB).  Each device holds its own copy until the process ends; host code
// reaches it only through the address scratch_on_device() looked up for that
// device (the name itself is a host-side stand-in).
constexpr long long kScratchDoubles = 1LL << 17;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on device `dev` (the current device), looked up
// once per device.
double *scratch_on_device(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw DeviceError{"device index out of range for the scratch table"};
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        double *p = nullptr;
        HIP_TRY(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = p;
    }
    return g_scratch_dev[dev];
}
}  // namespace

// Slots of c_pdev held by live contexts (a launch's descriptions go to its
// context's slot, so contexts on other streams never overwrite them).
namespace {
std::mutex g_pslot_mu;
bool g_pslot_used[kPSlots] = {};
int take_persist_slot() {
    std::lock_guard<std::mutex> lock(g_pslot_mu);
    for (int k = 0; k < kPSlots; ++k)
        if (!g_pslot_used[k]) { g_pslot_used[k] = true; return k; }
    return -1;
}
void give_persist_slot(int k) {
    std::lock_guard<std::mutex> lock(g_pslot_mu);
    g_pslot_used[k] = false;
}
}  // namespace

struct tis_ctx {
    int n_rep = 0;
    tis_options opt{};
    hipStream_t s_main = nullptr, s_hb = nullptr, s_mb = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join_mb = nullptr, ev_join_rb = nullptr;
    std::vector<void *> owned;
    Dev D{};
    std::vector<RepParams> rp;
    std::vector<int> n_st3;
    int nb_bonded = 0;
    int n_bonded_terms = 0;
    // the step's multi-body terms by k_multibody (bead forces summed in LDS):
    // slices per replica and dynamic LDS bytes; else k_force's step path
    bool mb_lds = false;
    int mb_slices = 1;
    size_t mb_lds_bytes = 0;
    int max_n = 1;                    // particles of the largest replica
    int dev_id = 0;
    std::vector<long long> stream_cap;
    long long stream_total = 0;
    bool graph_ok = false;
    // one captured graph per parity of the state buffers it starts from (a graph
    // of an odd number of steps ends on the other parity)
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_steps = 0;
    // state buffers: two of each, `par` the one holding the current state; every
    // step writes the other and the host swaps D's pointers after launching it
    double4 *buf_pos = nullptr, *buf_old = nullptr, *buf_vel = nullptr, *buf_fmb = nullptr;
    int par = 0;
    void set_parity(int p) {
        par = p;
        const size_t n = (size_t)D.n_tot;
        D.pos = buf_pos + p * n; D.pos_nx = buf_pos + (1 - p) * n;
        D.old = buf_old + p * n; D.old_nx = buf_old + (1 - p) * n;
        D.vel = buf_vel + p * n; D.vel_nx = buf_vel + (1 - p) * n;
        D.f_mb = buf_fmb + p * n; D.f_mb_nx = buf_fmb + (1 - p) * n;
    }
    void drop_graphs() {
        for (hipGraphExec_t &g : graph)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    }
    // launches of tis_step: graph replays, steps replayed from graphs, steps
    // launched one by one
    long long n_graph_launch = 0, n_graph_step = 0, n_eager_step = 0;
    // persistent step (k_persist): taken when set up at creation; launches and steps since
    bool persist = false, persist_ready = false;
    std::string persist_why = "not set up";
    int n_cu = 0, p_G = 0, p_q = 0;
    int *p_mem = nullptr;             // barrier counters, then status, then the XCC table
    int p_slot = -1;                  // slot of c_pdev (persist.h)
    size_t p_bar_ints = 0;
    long long n_persist_launch = 0, n_persist_step = 0;
    // kernel timing: when profiling is on, every launch of tis_step is bracketed
    // by HIP events on the stream it runs on (no graph), with the device idle
    // before and after it, and the durations are accumulated per kernel name
    bool profiling = false;
    struct Timed { const char *name; hipStream_t s; hipEvent_t a, b; };
    std::vector<Timed> pending;
    std::vector<hipEvent_t> event_pool;
    size_t event_next = 0;
    std::vector<std::string> knames;
    std::vector<double> ksum_ms;
    std::vector<long long> kcount;
    // replica exchange
    long long steps_done = 0;         // steps of tis_step since creation (exchange cadence)
    long long exchange_attempts = 0;
    bool all_same = false;            // every replica describes the same system
    // observables: sample cadence in steps of tis_step (0 = only tis_observe_now), contact cut
    long long obs_every = 0;
    double obs_cut = 7.2684;
    // force field as uploaded (for tis_set_force_field) and force-matching buffers
    std::vector<double> h_hbE, h_st3E;   // as uploaded at creation (sizes: all replicas' terms)
    std::vector<double2> h_st2rE;
    std::vector<std::vector<double>> type_R, type_eps, type_mass, type_q;
    double4 *fm_forces = nullptr;
    double *fm_t3 = nullptr, *fm_out3 = nullptr, *fm_small = nullptr;   // packed target / outputs, parameter staging
    double *fm_loss = nullptr, *fm_ion = nullptr, *fm_ghb = nullptr, *fm_gst3 = nullptr, *fm_gst2 = nullptr;
    double *scratch_dev = nullptr;   // this device's `scratch` (tis_force_match_noalloc)
    unsigned long long *hb_trace_buf = nullptr;   // tis_trace_hb ring (allocated when first switched on)

    ~tis_ctx() {
        if (p_slot >= 0) give_persist_slot(p_slot);
        drop_graphs();
        for (hipEvent_t e : event_pool) (void)hipEventDestroy(e);
        for (void *p : owned) (void)hipFree(p);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (ev_join_mb) (void)hipEventDestroy(ev_join_mb);
        if (ev_join_rb) (void)hipEventDestroy(ev_join_rb);
        if (s_main) (void)hipStreamDestroy(s_main);
        if (s_hb) (void)hipStreamDestroy(s_hb);
        if (s_mb) (void)hipStreamDestroy(s_mb);
    }

    int grid(int n) const { return (n + kBlock - 1) / kBlock; }
    // launch shape of the pair gather and the list build (tunable, results unchanged
    // up to the summation order of a particle's pair forces)
    int pair_block = kBlock, list_waves = kListWaves, hb_groups = 4;
    int pair_grid() const { return (int)(((long long)D.pair_lanes * D.n_tot + pair_block - 1) / pair_block); }
    int list_grid() const { return list_waves * 64 / kBlock; }

    hipEvent_t next_event() {
        if (event_next == event_pool.size()) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            event_pool.push_back(e);
        }
        return event_pool[event_next++];
    }
    // Brackets one launch with events when profiling.  The kernel runs alone:
    // with the streams overlapping, an event pair would also hold the time the
    // kernel's workgroups wait for CUs held by the other streams' kernels.
    template <class L>
    void timed(const char *name, hipStream_t s, L &&launch) {
        if (!profiling) { launch(); return; }
        hipEvent_t a = next_event(), b = next_event();
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipEventRecord(a, s));
        launch();
        HIP_TRY(hipEventRecord(b, s));
        HIP_TRY(hipDeviceSynchronize());
        pending.push_back({name, s, a, b});
    }
    void collect_times() {
        for (const Timed &t : pending) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
            size_t k = 0;
            while (k < knames.size() && knames[k] != t.name) ++k;
            if (k == knames.size()) { knames.push_back(t.name); ksum_ms.push_back(0); kcount.push_back(0); }
            ksum_ms[k] += ms;
            kcount[k] += 1;
        }
        pending.clear();
        event_next = 0;
    }

    // List rebuild of the replicas whose displacement trigger fired (the others
    // exit at once): cell sort and neighbour slots on `sl`, live HB candidates and
    // the resolution's site tables on `sh` (they depend on positions only).
    // The cell sort of the replicas whose trigger fired, over many workgroups.
    void launch_bin(hipStream_t sl) {
        const dim3 g((max_n + 1023) / 1024, n_rep);
        timed("k_bin_count", sl, [&] { k_bin_count<<<g, 1024, 0, sl>>>(D); });
        timed("k_bin_scan", sl, [&] { k_bin_scan<<<n_rep, 1024, 0, sl>>>(D); });
        timed("k_bin_place", sl, [&] { k_bin_place<<<g, 1024, 0, sl>>>(D); });
    }
    void launch_rebuild_lists(hipStream_t sl) {
        launch_bin(sl);
        timed("k_list", sl, [&] { k_list<<<list_grid(), kBlock, 0, sl>>>(D, 1); });
    }
    // lag: inside a step, beside the resolution, into the table buffer it is not
    // reading (switched by the step's close)
    void launch_rebuild_hb(hipStream_t sh, int lag) {
        timed("k_hbstatus", sh, [&] { k_hbstatus<<<dim3(max_hb_chunks(), n_rep), kStatusBlock, 0, sh>>>(D, lag); });
        timed("k_hbsites", sh, [&] { k_hbsites<<<n_rep, 1024, 2 * max_site_lds(), sh>>>(D, lag); });
    }
    void launch_rebuild() {
        launch_rebuild_lists(s_main);
        launch_rebuild_hb(s_main, 0);
    }
    // A step rebuilds at its start when the previous step's trigger fired; after
    // the last step of a call, do that rebuild now so that lists, HB statuses and
    // evaluations reflect the current positions, and clear the trigger.
    void finish_pending_rebuild() {
        launch_rebuild();
        k_clear_rebuild<<<grid(n_rep), kBlock, 0, s_main>>>(D);
        HIP_TRY(hipGetLastError());
    }
    int max_hb_chunks() const {
        int m = 1;
        for (const RepParams &p : rp) m = std::max(m, (p.nhb + kStatusBlock - 1) / kStatusBlock);
        return m;
    }
    size_t max_site_lds() const {
        int m = 1;
        for (const RepParams &p : rp) m = std::max(m, p.nsite);
        return (size_t)m * sizeof(int);
    }

    // HB terms on stream st: energies and gates, the valence resolution, forces.
    // The groups' energy sums and serial depths are reduced by k_hbsum, or in a
    // step (mode bit 0) by the step's close in k_integrate.
    void launch_hb(hipStream_t st, int mode) {
        launch_hb_energy(st);
        launch_hb_resolution(st, mode);
    }
    void launch_hb_energy(hipStream_t st) {
        const dim3 ge((kActMax + kBlock - 1) / kBlock, n_rep);
        timed("k_hbenergy", st, [&] {
            if (D.precision == TIS_PREC_FP64) k_hbenergy<double><<<ge, kBlock, 0, st>>>(D);
            else k_hbenergy<float><<<ge, kBlock, 0, st>>>(D);
        });
    }
    void launch_hb_resolution(hipStream_t st, int mode) {
        timed("k_hbres", st, [&] {
            if (D.precision == TIS_PREC_FP64) k_hbres<double><<<dim3(hb_groups, n_rep), kResBlock, kHbLds, st>>>(D, mode);
            else k_hbres<float><<<dim3(hb_groups, n_rep), kResBlock, kHbLds, st>>>(D, mode);
        });
        if (!(mode & 1)) k_hbsum<<<grid(n_rep), kBlock, 0, st>>>(D);
    }

    // Energies of every replica at its current configuration, with the pair
    // forces left untouched (eval mode 2); multi-body forces land in f_mb and are
    // cleared by the caller.  with_hb: include the HB terms (with their valence
    // resolution, which does not advance any stream).
    void launch_energies(bool with_hb) {
        k_zero_energies<<<grid(n_rep), kBlock, 0, s_main>>>(D);
        if (with_hb) launch_hb(s_main, 0);
        const int nmb = n_rep + nb_bonded, npr = pair_grid();
        const unsigned mb = TIS_TERM_STACK | TIS_TERM_BACKBONE, pr = TIS_TERM_SR | TIS_TERM_LR;
        if (D.precision == TIS_PREC_FP64) {
            k_force<double, 0><<<nmb, kBlock, 0, s_main>>>(D, nb_bonded, mb, kEvalFull);
            k_force<double, 1><<<npr, pair_block, 0, s_main>>>(D, nb_bonded, pr, kEvalEnergy);
        } else {
            k_force<float, 0><<<nmb, kBlock, 0, s_main>>>(D, nb_bonded, mb, kEvalFull);
            k_force<float, 1><<<npr, pair_block, 0, s_main>>>(D, nb_bonded, pr, kEvalEnergy);
        }
    }
    // One replica-exchange attempt, all on the device (the sequence documented
    // at tis_exchange_now in tis_gpu.h).
    void launch_exchange() {
        const int parity = (int)(exchange_attempts & 1);
        int max_n = 1;
        for (const RepParams &p : rp) max_n = std::max(max_n, p.n);
        const dim3 gs(grid(max_n), (n_rep + 1) / 2);
        launch_rebuild();                        // a trigger pending from the last step
        launch_energies(true);
        k_remd_collect<<<grid(n_rep), kBlock, 0, s_main>>>(D, 0);
        k_remd_mark<<<grid(n_rep), kBlock, 0, s_main>>>(D, parity);
        k_remd_swap<<<gs, kBlock, 0, s_main>>>(D, parity, 0);
        launch_rebuild();
        launch_energies(false);
        k_remd_collect<<<grid(n_rep), kBlock, 0, s_main>>>(D, 1);
        k_remd_decide<<<grid(n_rep), kBlock, 0, s_main>>>(D, exchange_attempts);
        k_remd_swap<<<gs, kBlock, 0, s_main>>>(D, parity, 1);
        launch_rebuild();
        k_clear_rebuild<<<grid(n_rep), kBlock, 0, s_main>>>(D);
        k_zero_fmb<<<grid(D.n_tot), kBlock, 0, s_main>>>(D);
        HIP_TRY(hipGetLastError());
        ++exchange_attempts;
    }

    // One sample of the observables of every replica at its current
    // configuration: a pending list rebuild is done now (the next step would
    // do the same one; its trigger is cleared), then every term's energy is
    // evaluated with the HB valence resolution (no stream advances, pair forces
    // untouched, multi-body forces cleared), then contacts, shape, energies and
    // HB / stack outcomes are added to the accumulators.
    void launch_observe() {
        k_obs_keep<<<n_rep, kBlock, 0, s_main>>>(D, 0);
        launch_rebuild();
        launch_energies(true);
        k_clear_rebuild<<<grid(n_rep), kBlock, 0, s_main>>>(D);
        k_zero_fmb<<<grid(D.n_tot), kBlock, 0, s_main>>>(D);
        k_obs_contacts<<<grid(D.n_tot), kBlock, 0, s_main>>>(D, obs_cut * obs_cut);
        k_obs_reduce<<<n_rep, kBlock, 0, s_main>>>(D);
        k_obs_keep<<<n_rep, kBlock, 0, s_main>>>(D, 1);   // the step's energies for the output files
        HIP_TRY(hipGetLastError());
    }
    void reset_observables() {
        int nhb = 0, nst3 = 0;
        for (const RepParams &p : rp) { nhb += p.nhb; nst3 += p.nst3; }
        HIP_TRY(hipMemsetAsync(D.obs_rep, 0, (size_t)kObsRep * n_rep * sizeof(double), s_main));
        HIP_TRY(hipMemsetAsync(D.obs_p, 0, (size_t)D.n_tot * sizeof(double4), s_main));
        HIP_TRY(hipMemsetAsync(D.obs_hb, 0, std::max(nhb, 1) * sizeof(double), s_main));
        HIP_TRY(hipMemsetAsync(D.obs_st3, 0, std::max(nst3, 1) * sizeof(double), s_main));
        HIP_TRY(hipMemsetAsync(D.obs_cnt, 0, 4 * (size_t)n_rep * sizeof(int), s_main));
    }

    // One Langevin step.  The HB resolution and the multi-body terms run on
    // their own streams, concurrently with the pair gather; the integrator joins all.
    // With D.hb_lag the HB tables of a rebuild are built on the multi-body stream,
    // beside the resolution, which reads the tables of the previous rebuild: every
    // candidate within its gate (r <= r0 + 2) is listed in both, because
    // r0 + 2 + (skin + the last step's displacement) < r_sr (checked at creation),
    // so the live set, and with it the resolution, is the same.
    void launch_step() {
        HIP_TRY(hipEventRecord(ev_fork, s_main));
        HIP_TRY(hipStreamWaitEvent(s_hb, ev_fork, 0));
        if (!D.hb_lag) launch_rebuild_hb(s_hb, 0);
        HIP_TRY(hipStreamWaitEvent(s_mb, ev_fork, 0));
        launch_hb_energy(s_hb);
        launch_hb_resolution(s_hb, 1);
        HIP_TRY(hipEventRecord(ev_join, s_hb));
        timed("k_multibody", s_mb, [&] {
            if (mb_lds) {
                const dim3 g(mb_slices, n_rep);
                if (D.precision == TIS_PREC_FP64) k_multibody<double><<<g, kMbBlock, mb_lds_bytes, s_mb>>>(D);
                else k_multibody<float><<<g, kMbBlock, mb_lds_bytes, s_mb>>>(D);
                return;
            }
            const int nmb = n_rep * D.st_slices + nb_bonded;
            if (D.precision == TIS_PREC_FP64) k_force<double, 0><<<nmb, kBlock, 0, s_mb>>>(D, nb_bonded, 31u, kEvalStep);
            else k_force<float, 0><<<nmb, kBlock, 0, s_mb>>>(D, nb_bonded, 31u, kEvalStep);
        });
        // the step's random forces (Philox draws), off the integrator's path
        if (D.rng == TIS_RNG_PHILOX) timed("k_noise", s_mb, [&] { k_noise<<<grid(D.n_tot), kBlock, 0, s_mb>>>(D); });
        HIP_TRY(hipEventRecord(ev_join_mb, s_mb));
        if (D.hb_lag) {   // after the multi-body terms on their stream (the shorter chain)
            launch_rebuild_hb(s_mb, 1);
            HIP_TRY(hipEventRecord(ev_join_rb, s_mb));
        }
        launch_rebuild_lists(s_main);
        timed("k_pairs", s_main, [&] {
            if (D.precision == TIS_PREC_FP64) k_force<double, 1><<<pair_grid(), pair_block, 0, s_main>>>(D, nb_bonded, 31u, kEvalStep);
            else k_force<float, 1><<<pair_grid(), pair_block, 0, s_main>>>(D, nb_bonded, 31u, kEvalStep);
        });
        HIP_TRY(hipStreamWaitEvent(s_main, ev_join, 0));
        // the close reads the new tables' counts (recorded after the multi-body terms on their stream)
        if (D.hb_lag) HIP_TRY(hipStreamWaitEvent(s_main, ev_join_rb, 0));
        else HIP_TRY(hipStreamWaitEvent(s_main, ev_join_mb, 0));
        timed("k_integrate", s_main, [&] { k_integrate<<<grid(D.n_tot), kBlock, 0, s_main>>>(D); });
        set_parity(par ^ 1);   // the new state is in the other buffers
    }

    // Force matching on stream st (tis_force_match, tis_force_match_noalloc):
    // forward, the step's forces with a fresh long-range term; residual v = F -
    // target into D.jv with its squared sum added to *loss; backward, force
    // tangents along v into D.jf and the parameter sums (ion radii into ion[0..2],
    // per-term values at global indices, or summed at replica-local ones when
    // local).  Multi-body forces are cleared at the end.
    void launch_force_match(hipStream_t st, const double *target, double4 *forces4, double *forces3, double *loss,
                            double *ion, double *ghb, double *gst3, double *gst2, int local) {
        const int N = D.n_tot;
        k_zero_fmb<<<grid(N), kBlock, 0, st>>>(D);
        launch_hb(st, 0);
        const int nmb = n_rep + nb_bonded, npr = pair_grid();
        if (D.precision == TIS_PREC_FP64) {
            k_force<double, 0><<<nmb, kBlock, 0, st>>>(D, nb_bonded, 31u, kEvalForce);
            k_force<double, 1><<<npr, pair_block, 0, st>>>(D, nb_bonded, 31u, kEvalForce);
        } else {
            k_force<float, 0><<<nmb, kBlock, 0, st>>>(D, nb_bonded, 31u, kEvalForce);
            k_force<float, 1><<<npr, pair_block, 0, st>>>(D, nb_bonded, 31u, kEvalForce);
        }
        k_fm_residual<<<grid(N), kBlock, 0, st>>>(D, target, forces4, forces3, loss);
        k_jvp_pairs<<<grid(kJvpLanes * N), kBlock, 0, st>>>(D, ion);
        k_jvp_multibody<<<nmb, kBlock, 0, st>>>(D, nb_bonded, gst3, gst2, local);
        k_jvp_hb<<<dim3((kActMax + kBlock - 1) / kBlock, n_rep), kBlock, 0, st>>>(D, ghb, local);
        k_zero_fmb<<<grid(N), kBlock, 0, st>>>(D);
        HIP_TRY(hipGetLastError());
    }

    RepState get_rs(int r) {
        RepState s;
        HIP_TRY(hipMemcpy(&s, D.rs + r, sizeof s, hipMemcpyDeviceToHost));
        return s;
    }
    void put_rs(int r, const RepState &s) {
        HIP_TRY(hipMemcpy(D.rs + r, &s, sizeof s, hipMemcpyHostToDevice));
    }
    std::vector<RepState> rs_host;    // n_rep, sized at creation (check_flags allocates nothing)
    int *p_status() const { return p_mem + p_bar_ints; }
    int *p_xcc() const { return p_mem + p_bar_ints + kPCtr; }
    // One cooperative launch of k_persist for n steps (n = 0: the placement
    // check only), from the current state buffers; the host's parity follows.
    void launch_persist(long long n) {
        PArgs A{p_slot, p_mem, p_status(), p_xcc(), (int)n, p_G, p_q, hb_groups};
        Dev Dv[2] = {D, D};
        std::swap(Dv[1].pos, Dv[1].pos_nx);
        std::swap(Dv[1].old, Dv[1].old_nx);
        std::swap(Dv[1].vel, Dv[1].vel_nx);
        std::swap(Dv[1].f_mb, Dv[1].f_mb_nx);
        HIP_TRY(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_pdev), Dv, sizeof Dv, 2 * sizeof(Dev) * p_slot,
                                       hipMemcpyHostToDevice, s_main));
        HIP_TRY(hipMemsetAsync(p_mem, 0, (p_bar_ints + kPCtr) * sizeof(int), s_main));
        void *args[] = {&A};
        const void *fn = D.precision == TIS_PREC_FP64 ? (const void *)k_persist<double> : (const void *)k_persist<float>;
        HIP_TRY(hipLaunchCooperativeKernel(fn, dim3(n_cu), dim3(kPBlock), args, kHbLds, s_main));
        set_parity(par ^ (int)(n & 1));
        if (n > 0) { n_persist_launch += 1; n_persist_step += n; }
    }
    // Status bits of the last persistent launch (synchronises the stream).
    int persist_status() {
        int st = 0;
        HIP_TRY(hipMemcpyAsync(&st, p_status(), sizeof st, hipMemcpyDeviceToHost, s_main));
        HIP_TRY(hipStreamSynchronize(s_main));
        return st;
    }
    void check_flags() {
        std::vector<RepState> &s = rs_host;
        HIP_TRY(hipMemcpy(s.data(), D.rs, n_rep * sizeof(RepState), hipMemcpyDeviceToHost));
        for (int r = 0; r < n_rep; ++r) {
            if (s[r].overflow & 1) throw Fail{TIS_EOVERFLOW, "neighbour slot overflow (more SR/LR partners than the slot set up at creation) in replica " + std::to_string(r)};
            if (s[r].overflow & 2) throw Fail{TIS_EOVERFLOW, "HB working set above the LDS caps in replica " + std::to_string(r)};
            if (s[r].overflow & 4) throw Fail{TIS_EOVERFLOW, "uniform stream exhausted in replica " + std::to_string(r)};
            if (s[r].overflow & 32)
                throw Fail{TIS_ENONFINITE, "non-finite coordinates in replica " + std::to_string(r)};
            if (s[r].overflow & 16)
                throw Fail{TIS_EOVERFLOW, "HB valence resolution: a site has more live candidates than the kernel holds "
                                          "(64) in replica " + std::to_string(r)};
        }
    }
};

namespace {

constexpr double kSigLJh = 1.5852;

// read_maxi_key derived constants (fast.cpp:2895-2944): friction and noise per
// type (gamma = 0.555 R), and the 9 x 9 pair tables (D_LJ = 3.2 between RNA
// beads, R_a + R_b otherwise; E_LJ = sqrt(eps_a eps_b); F_LJ = 12 E_LJ / 1.5852^2;
// Q_C lB).  P.T and P.h must be set.
void type_tables(RepParams &P, const double *R, const double *eps, const double *q, const double *mass, double lB) {
    for (int t = 0; t < kNT; ++t) {
        const double visc = 0.555 * R[t];
        P.mass[t] = mass[t];
        P.K1[t] = std::exp(-visc / mass[t] * P.h);
        P.K2[t] = (1.0 - P.K1[t]) / visc;
        P.sigma[t] = std::sqrt(2.0 * visc * P.T / P.h);
    }
    for (int a = 0; a < kNT; ++a)
        for (int b = 0; b < kNT; ++b) {
            const int lo = std::min(a, b), hi = std::max(a, b), ab = a * kNT + b;
            const double d = (lo <= 5 && hi <= 5) ? 3.2 : R[lo] + R[hi];
            P.dlj[ab] = d;
            P.d2lj[ab] = d * d;
            const double dm = d - 0.4406141861 * kSigLJh;
            P.d2min[ab] = dm * dm;
            P.elj[ab] = std::sqrt(eps[lo] * eps[hi]);
            P.flj[ab] = 12.0 * P.elj[ab] / (kSigLJh * kSigLJh);
            P.qlb[ab] = q[lo] * q[hi] * lB;
        }
}

// True when two replica descriptions are the same system: particle types and
// classes, bonds, angles, stacks and HB candidates (with their sites) equal.
// Temperatures, boxes, coordinates and T-dependent energies may differ.
bool same_system(const tis_system &a, const tis_system &b) {
    auto eq = [](const auto *x, const auto *y, long long n) {
        return n <= 0 || (x && y && std::memcmp(x, y, sizeof(*x) * (size_t)n) == 0);
    };
    return a.n_particles == b.n_particles && a.n_beads == b.n_beads && a.n_bonds == b.n_bonds &&
           a.n_angles == b.n_angles && a.n_st3 == b.n_st3 && a.n_st2 == b.n_st2 && a.n_hb == b.n_hb &&
           a.n_hb_sites == b.n_hb_sites && a.n_hb_incid == b.n_hb_incid && eq(a.type, b.type, a.n_particles) &&
           eq(a.cls, b.cls, a.n_particles) && eq(a.bond_ij, b.bond_ij, 2LL * a.n_bonds) &&
           eq(a.angle_ijk, b.angle_ijk, 3LL * a.n_angles) && eq(a.st3_beads, b.st3_beads, 6LL * a.n_st3) &&
           eq(a.st2_beads, b.st2_beads, 7LL * a.n_st2) && eq(a.hb_beads, b.hb_beads, 6LL * a.n_hb) &&
           eq(a.hb_site_idx, b.hb_site_idx, (long long)a.n_hb_incid);
}

// Validates one replica description; throws Fail on the first problem found.
void validate(const tis_system &s, int r) {
    auto bad = [&](const std::string &m) { throw Fail{TIS_EINVAL, "replica " + std::to_string(r) + ": " + m}; };
    if (s.n_particles <= 0 || s.n_beads <= 0 || s.n_beads > s.n_particles) bad("particle counts");
    if (!s.type || !s.cls || !s.pos || !s.pos_old || !s.vel) bad("missing particle arrays");
    if (s.recentre < 0 || s.recentre >= s.n_beads) bad("recentring bead out of range");
    for (int c = 0; c < 3; ++c)
        if (!(s.box[c] > 0)) bad("box");
    if (!(s.T > 0) || !(s.h > 0)) bad("T and h must be positive");
    for (int i = 0; i < s.n_particles; ++i) {
        if (s.type[i] < 0 || s.type[i] >= TIS_NTYPES) bad("type out of range");
        if (s.cls[i] < 0 || s.cls[i] > 5) bad("class out of range");
    }
    auto chk = [&](const int *a, long n, int hi, const char *what) {
        if (n > 0 && !a) bad(std::string("missing ") + what);
        for (long q = 0; q < n; ++q)
            if (a[q] < 0 || a[q] >= hi) bad(std::string(what) + " index out of range");
    };
    chk(s.bond_ij, 2L * s.n_bonds, s.n_beads, "bond");
    chk(s.angle_ijk, 3L * s.n_angles, s.n_beads, "angle");
    chk(s.st3_beads, 6L * s.n_st3, s.n_beads, "tertiary stack");
    chk(s.st2_beads, 7L * s.n_st2, s.n_beads, "secondary stack");
    chk(s.hb_beads, 6L * s.n_hb, s.n_beads, "HB");
    if (s.n_hb > 0) {
        if (!s.hb_site_off || !s.site_hb_off || !s.site_valence || !s.hb_geo || !s.hb_E || !s.hb_dode) bad("HB tables");
        if (s.hb_site_off[0] != 0 || s.hb_site_off[s.n_hb] != s.n_hb_incid) bad("HB incidence offsets");
        if (s.site_hb_off[0] != 0 || s.site_hb_off[s.n_hb_sites] != s.n_hb_incid) bad("site offsets");
        chk(s.hb_site_idx, s.n_hb_incid, s.n_hb_sites, "HB site");
        chk(s.site_hb_idx, s.n_hb_incid, s.n_hb, "site HB");
    }
    if (s.n_st2 > 4096) bad("more than 4096 secondary stacks");
    if (s.n_hb_incid >= (1 << 19)) bad("2^19 or more HB site incidences");
    for (int k = 0; k < 2 * s.n_st3; ++k)
        if (s.st3_disable && (s.st3_disable[k] < -1 || s.st3_disable[k] >= s.n_st2)) bad("st3_disable");
}

}  // namespace

// Launch-shape options: lanes per particle a power of two up to 64, pair block a
// multiple of 64 up to the kernels' launch bound (and of the lanes), list grid
// 64 .. 65536 wavefronts, 1 .. 16 workgroups per replica in the HB resolution.
static void check_launch_options(const tis_options &o) {
    const int l = o.pair_lanes, b = o.pair_block, w = o.list_waves;
    if (l < 1 || l > 64 || (l & (l - 1))) throw Fail{TIS_EINVAL, "pair_lanes must be a power of two, 1..64"};
    if (b < 64 || b > kBlock || b % 64) throw Fail{TIS_EINVAL, "pair_block must be a multiple of 64, at most 256"};
    if (w < 64 || w > 65536 || (w * 64) % kBlock) throw Fail{TIS_EINVAL, "list_waves must be 64..65536, a multiple of 4"};
    if (o.graph_steps < 0) throw Fail{TIS_EINVAL, "graph_steps"};
    if (o.hb_groups < 1 || o.hb_groups > kMaxHbGroups) throw Fail{TIS_EINVAL, "hb_groups must be 1..16"};
    if (o.hb_table_lag != 0 && o.hb_table_lag != 1) throw Fail{TIS_EINVAL, "hb_table_lag must be 0 or 1"};
}

extern "C" void tis_default_options(tis_options *o) {
    if (!o) return;
    o->precision = TIS_PREC_FP64;
    o->rng = TIS_RNG_PHILOX;
    o->seed = 1;
    o->noise_scale = 1.0;
    o->skin = 3.0;
    o->r_sr = 15.0;
    o->r_lr = 53.0;
    o->lr_every = 10;
    o->use_graph = 1;
    o->graph_steps = 50;
    o->exchange_every = 0;
    o->pair_lanes = kPairLanes;
    o->pair_block = kBlock;
    o->list_waves = kListWaves;
    o->hb_groups = 4;
    o->hb_table_lag = 1;
    o->persistent = 0;
}

static int fail(const Fail &f) {
    tis_host::set_error(f.msg);
    return f.code;
}
static int fail_dev(const DeviceError &e) {
    tis_host::set_error(e.msg);
    return TIS_EDEVICE;
}

// Runs an API body, mapping exceptions to the C error codes.
template <class F>
static int guarded(F &&body) {
    try {
        return body();
    } catch (const Fail &f) {
        return fail(f);
    } catch (const DeviceError &e) {
        return fail_dev(e);
    } catch (const std::bad_alloc &) {
        tis_host::set_error("out of host memory");
        return TIS_ENOMEM;
    }
}

// Sets up the persistent step (persist.h) when the context and the device allow
// it: Philox draws, HB tables built beside the resolution, multi-body forces
// summed in LDS, at least one replica per XCD and kHbLds of LDS for one
// workgroup per CU; the placement check then runs once (a launch of no steps).
// Otherwise persist_why says why and tis_step keeps to launch_step.
static void setup_persistent(tis_ctx *c) {
    const Dev &D = c->D;
    auto no = [&](const std::string &why) { c->persist_why = why; };
    if (!c->opt.persistent) return no("switched off (tis_options.persistent = 0)");
    if (D.rng != TIS_RNG_PHILOX) return no("reference stream draws");
    if (!D.hb_lag) return no("HB tables not built beside the resolution");
    if (!c->mb_lds) return no("multi-body forces not summed in LDS");
    int coop = 0;
    HIP_TRY(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, c->dev_id));
    HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c->dev_id));
    if (!coop) return no("no cooperative launch");
    if (c->n_cu % kXcds) return no("CU count not a multiple of 8");
    if (c->n_rep < kXcds) return no("fewer replicas than XCDs");
    c->p_q = (c->n_rep + kXcds - 1) / kXcds;
    c->p_G = c->n_cu / kXcds / c->p_q;
    if (c->p_G < c->hb_groups + 2) return no("too few CUs per replica for the HB groups and the other roles");
    const size_t sites_lds = 4 * ((size_t)kAIncMax + kASiteMax + 1 + kMaxHbChunks + 1 + 32) + 2 * c->max_site_lds();
    if (sites_lds > kHbLds) return no("HB sites above the LDS of the persistent workgroup");
    if (c->mb_lds_bytes + 4096 + 8 * (kPBlock / 64) > kHbLds) return no("multi-body window above the LDS");
    const void *fn = D.precision == TIS_PREC_FP64 ? (const void *)k_persist<double> : (const void *)k_persist<float>;
    HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kHbLds));
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kPBlock, kHbLds));
    if (per_cu < 1) return no("the persistent workgroup does not fit a CU");
    c->p_bar_ints = (3 * (size_t)c->n_rep + 1) * kPCtr;
    c->p_mem = dalloc<int>(c->owned, c->p_bar_ints + kPCtr + c->n_cu);
    c->p_slot = take_persist_slot();
    if (c->p_slot < 0) return no("all constant-memory slots of the persistent step in use");
    int st = 0;
    try {
        c->launch_persist(0);
        st = c->persist_status();
    } catch (const DeviceError &e) {
        (void)hipGetLastError();
        return no("placement check launch failed: " + e.msg);
    }
    if (st & PSTAT_PLACEMENT) return no("workgroups of one replica would not share an XCD");
    if (st) return no("placement check failed");
    c->persist_ready = c->persist = true;
    c->persist_why = "";
}

extern "C" int tis_create(const tis_system *reps, int n_rep, const tis_options *opt, tis_ctx **out) {
    if (!reps || n_rep <= 0 || !out) return fail(Fail{TIS_EINVAL, "tis_create: bad arguments"});
    *out = nullptr;
    tis_ctx *c = new (std::nothrow) tis_ctx();
    if (!c) return fail(Fail{TIS_ENOMEM, "tis_create: out of memory"});
    try {
        c->n_rep = n_rep;
        c->rs_host.resize(n_rep);
        if (opt) c->opt = *opt;
        else tis_default_options(&c->opt);
        const tis_options &o = c->opt;
        if (o.precision != TIS_PREC_FP64 && o.precision != TIS_PREC_MIXED) throw Fail{TIS_EINVAL, "precision"};
        if (o.rng != TIS_RNG_PHILOX && o.rng != TIS_RNG_STREAM) throw Fail{TIS_EINVAL, "rng"};
        if (!(o.r_sr > 0) || !(o.r_lr >= o.r_sr) || !(o.skin >= 0) || o.lr_every <= 0) throw Fail{TIS_EINVAL, "list settings"};
        if (o.exchange_every < 0) throw Fail{TIS_EINVAL, "exchange_every"};
        check_launch_options(o);
        for (int r = 0; r < n_rep; ++r) validate(reps[r], r);
        int dev_count = 0;
        HIP_TRY(hipGetDeviceCount(&dev_count));
        if (dev_count <= 0) throw Fail{TIS_EDEVICE, "no HIP device"};
        HIP_TRY(hipGetDevice(&c->dev_id));
        c->scratch_dev = scratch_on_device(c->dev_id);
        HIP_TRY(hipStreamCreateWithFlags(&c->s_main, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&c->s_hb, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&c->s_mb, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&c->ev_join_mb, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&c->ev_join_rb, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
        HIP_TRY(hipFuncSetAttribute((const void *)k_hbres<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kHbLds));
        HIP_TRY(hipFuncSetAttribute((const void *)k_hbres<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kHbLds));
        // ---- offsets and per-replica parameters
        int p0 = 0, b0 = 0, a0 = 0, s30 = 0, s20 = 0, h0 = 0, si0 = 0, i0 = 0, cell0 = 0;
        c->rp.resize(n_rep);
        for (int r = 0; r < n_rep; ++r) {
            const tis_system &s = reps[r];
            RepParams &P = c->rp[r];
            std::memset(&P, 0, sizeof P);
            P.p0 = p0; P.n = s.n_particles; P.nbeads = s.n_beads; P.kh = p0 + s.recentre;
            c->max_n = std::max(c->max_n, s.n_particles);
            for (int a = 0; a < 3; ++a) {
                P.L[a] = s.box[a];
                P.hL[a] = 0.5 * s.box[a];
                // cells at least r_lr / 2 wide (the list build searches +- 2 cells),
                // at most 20 per axis (the LDS counting sorts hold 2 kMaxCells)
                P.M[a] = std::min(20, std::max(1, (int)(2.0 * s.box[a] / o.r_lr)));
                P.lc[a] = s.box[a] / P.M[a];
            }
            if ((long)P.M[0] * P.M[1] * P.M[2] > kMaxCells) throw Fail{TIS_EINVAL, "box too large for the cell grid"};
            P.sr_h = (P.lc[0] >= o.r_sr && P.lc[1] >= o.r_sr && P.lc[2] >= o.r_sr) ? 1 : 2;
            P.T = s.T; P.lB = s.lB; P.h = s.h; P.vcap = 300.0 * s.T;
            P.rsr2 = o.r_sr * o.r_sr; P.rlr2 = o.r_lr * o.r_lr; P.skin = o.skin;
            P.lr_every = o.lr_every;
            P.cell0 = cell0;
            P.bond0 = b0; P.nbond = s.n_bonds; P.ang0 = a0; P.nang = s.n_angles;
            P.st30 = s30; P.nst3 = s.n_st3; P.st20 = s20; P.nst2 = s.n_st2;
            P.hb0 = h0; P.nhb = s.n_hb; P.site0 = si0; P.nsite = s.n_hb_sites; P.inc0 = i0; P.ninc = s.n_hb_incid;
            type_tables(P, s.type_R, s.type_eps, s.type_q, s.type_mass, s.lB);
            p0 += s.n_particles; b0 += s.n_bonds; a0 += s.n_angles; s30 += s.n_st3; s20 += s.n_st2;
            h0 += s.n_hb; si0 += s.n_hb_sites; i0 += s.n_hb_incid;
            cell0 += 2 * P.M[0] * P.M[1] * P.M[2] + 1;   // charged and uncharged parts of the cell order
        }
        const int N = p0;
        Dev &D = c->D;
        D.n_rep = n_rep; D.n_tot = N; D.precision = o.precision; D.rng = o.rng; D.seed = o.seed;
        D.noise_scale = o.noise_scale;
        D.pair_lanes = o.pair_lanes;
        c->pair_block = o.pair_block;
        c->list_waves = o.list_waves;
        c->hb_groups = o.hb_groups;
        D.hb_groups = o.hb_groups;
        auto &own = c->owned;
        D.rp = dalloc<RepParams>(own, n_rep);
        D.rs = dalloc<RepState>(own, n_rep);
        c->buf_pos = dalloc<double4>(own, 2 * (size_t)N);
        c->buf_old = dalloc<double4>(own, 2 * (size_t)N);
        c->buf_vel = dalloc<double4>(own, 2 * (size_t)N);
        c->buf_fmb = dalloc<double4>(own, 2 * (size_t)N);
        c->set_parity(0);
        D.f_sr = dalloc<double4>(own, N);
        D.nz = dalloc<double4>(own, N);
        D.maxd2 = dalloc<unsigned long long>(own, (size_t)n_rep * kMaxSlots * kSlotStride);
        D.close_cnt = dalloc<int>(own, (size_t)n_rep * kCloseStride);
        D.f_lr = dalloc<double4>(own, N);
        D.ptype = dalloc<int>(own, N);
        D.prep = dalloc<int>(own, N);
        D.cell_of = dalloc<int>(own, N);
        D.cell_rank = dalloc<int>(own, N);
        D.cell_start = dalloc<int>(own, cell0);
        D.cell_cnt = dalloc<int>(own, cell0);
        D.sorted = dalloc<int>(own, N);
        D.cpos = dalloc<double4>(own, N);
        D.crank = dalloc<int>(own, N);
        D.nb_start = dalloc<int>(own, N);
        D.nb_nsr = dalloc<int>(own, N);
        D.nb_nlr = dalloc<int>(own, N);
        D.nb_meta = dalloc<int4>(own, N);
        // ---- topology, converted to global indices
        std::vector<int> ptype(N), prep(N);
        std::vector<double4> pos(N), old(N), vel(N);
        std::vector<int2> bij;
        std::vector<double2> bkr, akt, st2rE;
        std::vector<int4> aijk;
        std::vector<int> st3b, st2b, hbb, hbd, hio, hi, so, sh, sv;
        std::vector<double> st3g, st3E, hbg, hbE;
        std::vector<int2> st3d;
        hio.push_back(0);
        so.push_back(0);
        int inc_base = 0;
        for (int r = 0; r < n_rep; ++r) {
            const tis_system &s = reps[r];
            const RepParams &P = c->rp[r];
            const int o0 = P.p0;
            for (int i = 0; i < s.n_particles; ++i) {
                const int charged = (s.cls[i] != 1 && s.cls[i] != 2) ? 1 : 0;
                ptype[o0 + i] = s.type[i] | (s.cls[i] << 4) | (charged << 8);
                prep[o0 + i] = r;
                pos[o0 + i] = make_double4(s.pos[3 * i], s.pos[3 * i + 1], s.pos[3 * i + 2], 0);
                old[o0 + i] = make_double4(s.pos_old[3 * i], s.pos_old[3 * i + 1], s.pos_old[3 * i + 2], 0);
                vel[o0 + i] = make_double4(s.vel[3 * i], s.vel[3 * i + 1], s.vel[3 * i + 2], 0);
            }
            for (int k = 0; k < s.n_bonds; ++k) {
                bij.push_back(make_int2(o0 + s.bond_ij[2 * k], o0 + s.bond_ij[2 * k + 1]));
                bkr.push_back(make_double2(s.bond_kr[2 * k], s.bond_kr[2 * k + 1]));
            }
            for (int k = 0; k < s.n_angles; ++k) {
                aijk.push_back(make_int4(o0 + s.angle_ijk[3 * k], o0 + s.angle_ijk[3 * k + 1],
                                         o0 + s.angle_ijk[3 * k + 2], 0));
                akt.push_back(make_double2(s.angle_kt[2 * k], s.angle_kt[2 * k + 1]));
            }
            for (int k = 0; k < s.n_st3; ++k) {
                for (int q = 0; q < 6; ++q) {
                    st3b.push_back(o0 + s.st3_beads[6 * k + q]);
                    st3g.push_back(s.st3_geo[6 * k + q]);
                }
                st3E.push_back(s.st3_E[k]);
                int d0 = s.st3_disable ? s.st3_disable[2 * k] : -1, d1 = s.st3_disable ? s.st3_disable[2 * k + 1] : -1;
                st3d.push_back(make_int2(d0, d1));
            }
            for (int k = 0; k < s.n_st2; ++k) {
                for (int q = 0; q < 7; ++q) st2b.push_back(o0 + s.st2_beads[7 * k + q]);
                st2b.push_back(0);
                st2rE.push_back(make_double2(s.st2_r0[k], s.st2_E[k]));
            }
            for (int k = 0; k < s.n_hb; ++k) {
                for (int q = 0; q < 6; ++q) {
                    hbb.push_back(o0 + s.hb_beads[6 * k + q]);
                    hbg.push_back(s.hb_geo[6 * k + q]);
                }
                hbE.push_back(s.hb_E[k]);
                hbd.push_back(s.hb_dode[k]);
                for (int q = s.hb_site_off[k]; q < s.hb_site_off[k + 1]; ++q) hi.push_back(P.site0 + s.hb_site_idx[q]);
                hio.push_back((int)hi.size());
            }
            for (int q = 0; q < s.n_hb_sites; ++q) {
                for (int e = s.site_hb_off[q]; e < s.site_hb_off[q + 1]; ++e) sh.push_back(P.hb0 + s.site_hb_idx[e]);
                so.push_back((int)sh.size());
                sv.push_back(s.site_valence[q]);
            }
            inc_base += s.n_hb_incid;
        }
        if (2 * c->max_site_lds() > 100 * 1024) throw Fail{TIS_EINVAL, "too many HB sites per replica"};
        HIP_TRY(hipFuncSetAttribute((const void *)k_hbsites, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(2 * c->max_site_lds())));
        // bonded / 1-3 exclusions: every bond (i, k) and angle end pair (i, k) has its
        // WCA removed once per term; the step path leaves the pair's WCA out of the
        // pair sum and removes it again only for a second or later term of the same pair
        {
            std::vector<int> ex((size_t)N * kExcl, -1);
            std::vector<unsigned char> bs(bij.size() + aijk.size(), 0);
            std::map<std::pair<int, int>, int> seen;
            auto add = [&](int i, int k, size_t term) {
                const std::pair<int, int> key(std::min(i, k), std::max(i, k));
                if (seen[key]++ > 0) { bs[term] = 1; return; }
                for (int a : {i, k}) {
                    const int b = a == i ? k : i;
                    int e = 0;
                    while (e < kExcl && ex[(size_t)a * kExcl + e] >= 0) ++e;
                    if (e == kExcl) throw Fail{TIS_EINVAL, "a particle has more than 8 bonded / 1-3 partners"};
                    ex[(size_t)a * kExcl + e] = b;
                }
            };
            for (size_t q = 0; q < bij.size(); ++q) add(bij[q].x, bij[q].y, q);
            for (size_t q = 0; q < aijk.size(); ++q) add(aijk[q].x, aijk[q].z, bij.size() + q);
            D.excl = dalloc<int>(own, ex.size()); upload(D.excl, ex);
            D.bsub = dalloc<unsigned char>(own, std::max<size_t>(bs.size(), 1)); upload(D.bsub, bs);
        }
        c->n_bonded_terms = (int)bij.size() + (int)aijk.size();
        c->nb_bonded = c->grid(c->n_bonded_terms);
        D.bond_ij = dalloc<int2>(own, bij.size()); upload(D.bond_ij, bij);
        D.bond_kr = dalloc<double2>(own, bkr.size()); upload(D.bond_kr, bkr);
        D.ang_ijk = dalloc<int4>(own, aijk.size()); upload(D.ang_ijk, aijk);
        D.ang_kt = dalloc<double2>(own, akt.size()); upload(D.ang_kt, akt);
        D.st3_b = dalloc<int>(own, st3b.size()); upload(D.st3_b, st3b);
        D.st3_g = dalloc<double>(own, st3g.size()); upload(D.st3_g, st3g);
        D.st3_E = dalloc<double>(own, st3E.size()); upload(D.st3_E, st3E);
        c->h_st3E = st3E;
        D.st3_dis = dalloc<int2>(own, st3d.size()); upload(D.st3_dis, st3d);
        D.st2_b = dalloc<int>(own, st2b.size()); upload(D.st2_b, st2b);
        D.st2_rE = dalloc<double2>(own, st2rE.size()); upload(D.st2_rE, st2rE);
        c->h_st2rE = st2rE;
        D.hb_b = dalloc<int>(own, hbb.size()); upload(D.hb_b, hbb);
        D.hb_g = dalloc<double>(own, hbg.size()); upload(D.hb_g, hbg);
        {   // HB tables rebuilt beside the resolution only where no candidate in its
            // gate can be missing from the previous rebuild's list: r0 + 2 + skin
            // plus 1 A for the displacement of the step that fires the trigger
            double r0max = 0.0;
            for (size_t k = 0; k < hbg.size(); k += 6) r0max = std::max(r0max, hbg[k]);
            D.hb_lag = o.hb_table_lag && r0max + 2.0 + o.skin + 1.0 < o.r_sr;
        }
        D.hb_E = dalloc<double>(own, hbE.size()); upload(D.hb_E, hbE);
        c->h_hbE = hbE;
        for (int r = 0; r < n_rep; ++r) {
            c->type_R.emplace_back(reps[r].type_R, reps[r].type_R + kNT);
            c->type_eps.emplace_back(reps[r].type_eps, reps[r].type_eps + kNT);
            c->type_mass.emplace_back(reps[r].type_mass, reps[r].type_mass + kNT);
            c->type_q.emplace_back(reps[r].type_q, reps[r].type_q + kNT);
        }
        D.hb_dode = dalloc<int>(own, hbd.size()); upload(D.hb_dode, hbd);
        D.hb_inc_off = dalloc<int>(own, hio.size()); upload(D.hb_inc_off, hio);
        D.hb_inc = dalloc<int>(own, hi.size()); upload(D.hb_inc, hi);
        D.site_off = dalloc<int>(own, so.size()); upload(D.site_off, so);
        D.site_hb = dalloc<int>(own, sh.size()); upload(D.site_hb, sh);
        D.site_val = dalloc<int>(own, sv.size()); upload(D.site_val, sv);
        // rank of every HB incidence within its site's static list (HB_ATOM order);
        // the j-th incidence of a candidate at a site takes the j-th occurrence of
        // that candidate in the site's list
        std::vector<int> rank(hi.size(), -1);
        for (size_t site = 0; site + 1 < so.size(); ++site) {
            for (int e = so[site]; e < so[site + 1]; ++e) {
                const int k = sh[e];
                int t = hio[k];
                while (t < hio[k + 1] && !(hi[t] == (int)site && rank[t] < 0)) ++t;
                if (t == hio[k + 1]) throw Fail{TIS_EINVAL, "HB site tables are not consistent"};
                rank[t] = e - so[site];
            }
        }
        for (int v : rank)
            if (v < 0 || v > 0x7fff) throw Fail{TIS_EINVAL, "HB site tables are not consistent"};
        D.inc_rank = dalloc<int>(own, rank.size()); upload(D.inc_rank, rank);
        D.chunk_cnt = dalloc<int>(own, (size_t)kMaxHbChunks * n_rep);
        for (const RepParams &p : c->rp)
            if (p.nhb > kMaxHbChunks * kStatusBlock) throw Fail{TIS_EINVAL, "more than 65536 HB candidates in a replica"};
        D.site_cnt = dalloc<int>(own, sv.size());
        HIP_TRY(hipMemset(D.site_cnt, 0, sv.size() * sizeof(int)));
        const size_t nhb_tot = hbE.size(), nsite_tot = sv.size();
        D.hb_status = dalloc<signed char>(own, nhb_tot);
        D.hb_out = dalloc<signed char>(own, nhb_tot);
        D.hb_en = dalloc<double>(own, (size_t)kActMax * n_rep);
        D.hb_gate = dalloc<unsigned char>(own, (size_t)kActMax * n_rep);
        D.hb_frc = dalloc<double>(own, 18 * (size_t)kActMax * n_rep);
        D.hb_iter_part = dalloc<int>(own, (size_t)kMaxHbGroups * n_rep);
        D.hb_e_part = dalloc<double>(own, 2 * (size_t)kMaxHbGroups * n_rep);
        D.tab_hb = (long long)nhb_tot + n_rep;
        D.tab_site = (long long)nsite_tot + n_rep;
        D.tab_inc = std::max<long long>(1, inc_base);
        D.act_hb = dalloc<int>(own, 2 * (size_t)D.tab_hb);
        D.act_tmp = dalloc<int>(own, nhb_tot);
        D.act_flag = dalloc<unsigned char>(own, nhb_tot);
        D.asite_val = dalloc<int>(own, 2 * (size_t)D.tab_site);
        D.asite_off = dalloc<int>(own, 2 * (size_t)D.tab_site);
        D.asite_list = dalloc<int>(own, 2 * (size_t)D.tab_inc);
        D.ahb_inc_off = dalloc<int>(own, 2 * (size_t)D.tab_hb);
        D.ahb_inc = dalloc<int>(own, 2 * (size_t)D.tab_inc);
        D.act_rec = dalloc<int4>(own, 4 * (size_t)D.tab_hb);
        D.st_energy = dalloc<double>(own, st3E.size());
        D.st_part = dalloc<double>(own, 2 * (size_t)kMaxStSlices * n_rep);
        {   // one stack per thread in a step: slices of kBlock stack slots (the
            // tertiary stacks padded to whole wavefronts)
            int mx = 1;
            for (const RepParams &q : c->rp) mx = std::max(mx, ((q.nst3 + 63) & ~63) + q.nst2);
            D.st_slices = std::min(16, (mx + kBlock - 1) / kBlock);
        }
        {   // k_multibody: each replica's multi-body terms by type, sorted by their
            // lowest bead and cut into slices of the replica's beads (one slice per
            // 256 terms of the largest replica), with the bead window each slice's
            // local terms reach; taken when the largest window fits the LDS
            int max_terms = 1;
            for (const RepParams &q : c->rp) max_terms = std::max(max_terms, q.nst3 + q.nst2 + q.nbond + q.nang);
            const int S = std::min(kMaxStSlices, std::max(1, (max_terms + 255) / 256)), nbond_tot = (int)bij.size();
            std::vector<int> terms, offs((size_t)n_rep * (S + 1) * 4), win(2 * (size_t)n_rep * S);
            int max_win = 1;
            for (int r = 0; r < n_rep; ++r) {
                const RepParams &P = c->rp[r];
                // (lowest bead, highest bead, term id) per type
                std::vector<std::array<int, 3>> lists[4];
                auto span = [](const int *b, int n) {
                    return std::make_pair(*std::min_element(b, b + n), *std::max_element(b, b + n));
                };
                for (int k = 0; k < P.nst3; ++k) {
                    const int g = P.st30 + k;
                    const auto m = span(&st3b[6 * (size_t)g], 6);
                    lists[0].push_back({m.first, m.second, g});
                }
                for (int k = 0; k < P.nst2; ++k) {
                    const int g = P.st20 + k;
                    const auto m = span(&st2b[8 * (size_t)g], 7);
                    lists[1].push_back({m.first, m.second, g});
                }
                for (int k = 0; k < P.nbond; ++k) {
                    const int g = P.bond0 + k, b2[2] = {bij[g].x, bij[g].y};
                    const auto m = span(b2, 2);
                    lists[2].push_back({m.first, m.second, g});
                }
                for (int k = 0; k < P.nang; ++k) {
                    const int g = P.ang0 + k, b3[3] = {aijk[g].x, aijk[g].y, aijk[g].z};
                    const auto m = span(b3, 3);
                    lists[3].push_back({m.first, m.second, nbond_tot + g});
                }
                for (int q = 0; q < S; ++q) {
                    win[2 * ((size_t)r * S + q)] = P.p0 + (int)((long long)q * P.nbeads / S);
                    win[2 * ((size_t)r * S + q) + 1] = P.p0 + (int)((long long)(q + 1) * P.nbeads / S);
                }
                for (int t = 0; t < 4; ++t) {
                    std::stable_sort(lists[t].begin(), lists[t].end(),
                                     [](const auto &a, const auto &b) { return a[0] < b[0]; });
                    const int base = (int)terms.size();
                    for (const auto &e : lists[t]) terms.push_back(e[2]);
                    for (int q = 0; q <= S; ++q) {
                        const int bead = P.p0 + (int)((long long)q * P.nbeads / S);
                        const auto it = q == S ? lists[t].end()
                                               : std::lower_bound(lists[t].begin(), lists[t].end(), bead,
                                                                  [](const auto &e, int v) { return e[0] < v; });
                        offs[((size_t)r * (S + 1) + q) * 4 + t] = base + (int)(it - lists[t].begin());
                    }
                    if (t == 0) continue;   // tertiary stacks add to f_mb directly
                    for (int q = 0; q < S; ++q) {   // window end: the highest bead of the slice's terms
                        const int a = offs[((size_t)r * (S + 1) + q) * 4 + t] - base;
                        const int b = offs[((size_t)r * (S + 1) + q + 1) * 4 + t] - base;
                        int &end = win[2 * ((size_t)r * S + q) + 1];
                        for (int e = a; e < b; ++e) end = std::max(end, lists[t][e][1] + 1);
                    }
                }
                for (int q = 0; q < S; ++q)
                    max_win = std::max(max_win, win[2 * ((size_t)r * S + q) + 1] - win[2 * ((size_t)r * S + q)]);
            }
            c->mb_slices = S;
            c->mb_lds_bytes = 3 * (size_t)max_win * sizeof(double);
            c->mb_lds = c->mb_lds_bytes <= 120 * 1024;
            if (c->mb_lds) {
                D.mb_terms = dalloc<int>(own, terms.size()); upload(D.mb_terms, terms);
                D.mb_off = dalloc<int>(own, offs.size()); upload(D.mb_off, offs);
                D.mb_win = dalloc<int>(own, win.size()); upload(D.mb_win, win);
                D.st_slices = S;
                HIP_TRY(hipFuncSetAttribute((const void *)k_multibody<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)c->mb_lds_bytes));
                HIP_TRY(hipFuncSetAttribute((const void *)k_multibody<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)c->mb_lds_bytes));
            }
        }
        D.maxwell = dalloc<double>(own, 3000 * (size_t)n_rep);
        D.stream_base = dalloc<long long>(own, n_rep);
        D.stream_len = dalloc<long long>(own, n_rep);
        c->stream_cap.assign(n_rep, 0);
        // ---- replica exchange: neighbours describing the same system may swap
        {
            std::vector<int> ok(n_rep, 0), walker(n_rep);
            for (int r = 0; r < n_rep; ++r) walker[r] = r;
            for (int r = 0; r + 1 < n_rep; ++r) ok[r] = same_system(reps[r], reps[r + 1]) ? 1 : 0;
            c->all_same = true;
            for (int r = 0; r + 1 < n_rep; ++r) c->all_same = c->all_same && ok[r];
            D.remd_ok = dalloc<int>(own, n_rep); upload(D.remd_ok, ok);
            D.walker = dalloc<int>(own, n_rep); upload(D.walker, walker);
            D.remd_state = dalloc<int>(own, n_rep);
            HIP_TRY(hipMemset(D.remd_state, 0, n_rep * sizeof(int)));
            D.remd_e = dalloc<double>(own, 4 * (size_t)n_rep);
            HIP_TRY(hipMemset(D.remd_e, 0, 4 * n_rep * sizeof(double)));
            D.remd_d = dalloc<double>(own, n_rep);
            HIP_TRY(hipMemset(D.remd_d, 0, n_rep * sizeof(double)));
            D.remd_stats = dalloc<long long>(own, 2 * (size_t)n_rep);
            HIP_TRY(hipMemset(D.remd_stats, 0, 2 * n_rep * sizeof(long long)));
        }
        // ---- observables (zeroed by dalloc)
        D.obs_rep = dalloc<double>(own, (size_t)kObsRep * n_rep);
        D.obs_p = dalloc<double4>(own, N);
        D.obs_hb = dalloc<double>(own, std::max<size_t>(1, h0));
        D.obs_st3 = dalloc<double>(own, std::max<size_t>(1, s30));
        D.obs_cnt = dalloc<int>(own, 4 * (size_t)n_rep);
        D.obs_keep = dalloc<double>(own, (size_t)kObsKeep * n_rep + std::max<size_t>(1, s30));
        D.obs_keep_hb = dalloc<signed char>(own, std::max<size_t>(1, h0));
        // ---- force matching: tangents, target, outputs
        D.jv = dalloc<double4>(own, N);
        D.jf = dalloc<double4>(own, N);
        c->fm_forces = dalloc<double4>(own, N);
        c->fm_t3 = dalloc<double>(own, 3 * (size_t)N);
        c->fm_out3 = dalloc<double>(own, 3 * (size_t)N);
        c->fm_small = dalloc<double>(own, std::max<size_t>({1, c->h_hbE.size(), c->h_st3E.size(), c->h_st2rE.size()}));
        c->fm_loss = dalloc<double>(own, 1);
        c->fm_ion = dalloc<double>(own, 3);
        c->fm_ghb = dalloc<double>(own, std::max<size_t>(1, c->h_hbE.size()));
        c->fm_gst3 = dalloc<double>(own, std::max<size_t>(1, c->h_st3E.size()));
        c->fm_gst2 = dalloc<double>(own, std::max<size_t>(1, c->h_st2rE.size()));
        upload(D.ptype, ptype);
        upload(D.prep, prep);
        upload(D.pos, pos);
        upload(D.old, old);
        upload(D.vel, vel);
        // ---- size the neighbour pool from a counting pass, then build
        std::vector<RepState> rs(n_rep);
        for (int r = 0; r < n_rep; ++r) {
            std::memset(&rs[r], 0, sizeof(RepState));
            rs[r].klok = 1;
            rs[r].need_rebuild = 1;
            rs[r].lr_due = 1;
        }
        upload(D.rs, rs);
        HIP_TRY(hipMemcpy(D.rp, c->rp.data(), n_rep * sizeof(RepParams), hipMemcpyHostToDevice));
        c->launch_bin(c->s_main);
        k_list<<<c->list_grid(), kBlock, 0, c->s_main>>>(D, 0);
        HIP_TRY(hipStreamSynchronize(c->s_main));
        // fixed per-particle slots: twice the largest count at set-up plus margin
        // (compaction under added salt raises the LR counts of the RNA core)
        long long pool_total = 0;
        for (int r = 0; r < n_rep; ++r) {
            RepState s = c->get_rs(r);
            RepParams &P = c->rp[r];
            P.cap_sr = (2 * s.max_sr + 64 + 15) / 16 * 16;
            P.cap_lr = (2 * s.max_lr + 128 + 15) / 16 * 16;
            const long long stride = (long long)P.cap_sr + P.cap_lr;
            if (stride * P.n > 0x7fffffffLL) throw Fail{TIS_EINVAL, "neighbour slots of one replica exceed 2^31 entries"};
            P.pool0 = pool_total;
            P.pool_cap = stride * P.n;
            pool_total += P.pool_cap;
        }
        D.pool = dalloc<int>(own, (size_t)pool_total);
        HIP_TRY(hipMemcpy(D.rp, c->rp.data(), n_rep * sizeof(RepParams), hipMemcpyHostToDevice));
        c->launch_rebuild();
        HIP_TRY(hipStreamSynchronize(c->s_main));
        for (int r = 0; r < n_rep; ++r) {
            RepState s = c->get_rs(r);
            s.need_rebuild = 0;
            c->put_rs(r, s);
        }
        c->check_flags();
        setup_persistent(c);
        HIP_TRY(hipGetLastError());
        *out = c;
        return TIS_OK;
    } catch (const Fail &f) {
        delete c;
        return fail(f);
    } catch (const DeviceError &e) {
        delete c;
        return fail_dev(e);
    } catch (const std::bad_alloc &) {
        delete c;
        tis_host::set_error("out of host memory");
        return TIS_ENOMEM;
    }
}

extern "C" void tis_destroy(tis_ctx *c) { delete c; }

extern "C" int tis_set_state(tis_ctx *c, int rep, const double *pos, const double *pos_old, const double *vel,
                             long long klok) {
    if (!c || rep < 0 || rep >= (c ? c->n_rep : 0) || !pos || !pos_old || !vel)
        return fail(Fail{TIS_EINVAL, "tis_set_state: bad arguments"});
    return guarded([&]() -> int {
        const RepParams &P = c->rp[rep];
        std::vector<double4> a(P.n), b(P.n), v(P.n);
        for (int i = 0; i < 3 * P.n; ++i)
            if (!std::isfinite(pos[i]) || !std::isfinite(pos_old[i]) || !std::isfinite(vel[i]))
                throw Fail{TIS_EINVAL, "tis_set_state: non-finite coordinate or velocity"};
        for (int i = 0; i < P.n; ++i) {
            a[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0);
            b[i] = make_double4(pos_old[3 * i], pos_old[3 * i + 1], pos_old[3 * i + 2], 0);
            v[i] = make_double4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0);
        }
        HIP_TRY(hipStreamSynchronize(c->s_main));
        HIP_TRY(hipMemcpy(c->D.pos + P.p0, a.data(), P.n * sizeof(double4), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(c->D.old + P.p0, b.data(), P.n * sizeof(double4), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(c->D.vel + P.p0, v.data(), P.n * sizeof(double4), hipMemcpyHostToDevice));
        std::vector<RepState> all(c->n_rep);
        HIP_TRY(hipMemcpy(all.data(), c->D.rs, c->n_rep * sizeof(RepState), hipMemcpyDeviceToHost));
        for (int r = 0; r < c->n_rep; ++r) all[r].need_rebuild = (r == rep);
        all[rep].klok = klok;
        all[rep].progress = 0;
        all[rep].lr_due = 1;
        all[rep].overflow = 0;
        HIP_TRY(hipMemcpy(c->D.rs, all.data(), c->n_rep * sizeof(RepState), hipMemcpyHostToDevice));
        HIP_TRY(hipMemsetAsync(c->D.maxd2 + (size_t)rep * kMaxSlots * kSlotStride, 0,
                               (size_t)kMaxSlots * kSlotStride * sizeof(unsigned long long), c->s_main));
        c->launch_rebuild();
        HIP_TRY(hipStreamSynchronize(c->s_main));
        HIP_TRY(hipMemcpy(all.data(), c->D.rs, c->n_rep * sizeof(RepState), hipMemcpyDeviceToHost));
        all[rep].need_rebuild = 0;
        HIP_TRY(hipMemcpy(c->D.rs, all.data(), c->n_rep * sizeof(RepState), hipMemcpyHostToDevice));
        c->check_flags();
        return TIS_OK;
    });
}

extern "C" int tis_get_state(tis_ctx *c, int rep, double *pos, double *pos_old, double *vel) {
    if (!c || rep < 0 || rep >= c->n_rep) return fail(Fail{TIS_EINVAL, "tis_get_state: bad arguments"});
    return guarded([&]() -> int {
        const RepParams &P = c->rp[rep];
        HIP_TRY(hipStreamSynchronize(c->s_main));
        std::vector<double4> t(P.n);
        double4 *src[3] = {c->D.pos, c->D.old, c->D.vel};
        double *dst[3] = {pos, pos_old, vel};
        for (int a = 0; a < 3; ++a) {
            if (!dst[a]) continue;
            HIP_TRY(hipMemcpy(t.data(), src[a] + P.p0, P.n * sizeof(double4), hipMemcpyDeviceToHost));
            for (int i = 0; i < P.n; ++i) {
                dst[a][3 * i] = t[i].x;
                dst[a][3 * i + 1] = t[i].y;
                dst[a][3 * i + 2] = t[i].z;
            }
        }
        return TIS_OK;
    });
}

extern "C" int tis_set_step_counter(tis_ctx *c, int rep, long long klok) {
    if (!c || rep < 0 || rep >= c->n_rep || klok < 1) return fail(Fail{TIS_EINVAL, "tis_set_step_counter: bad arguments"});
    return guarded([&]() -> int {
        HIP_TRY(hipStreamSynchronize(c->s_main));
        RepState s = c->get_rs(rep);
        s.klok = klok;
        s.lr_due = (klok % c->rp[rep].lr_every) == 1 || s.progress == 0;
        c->put_rs(rep, s);
        return TIS_OK;
    });
}

extern "C" int tis_set_rng_epoch(tis_ctx *c, int rep, unsigned epoch) {
    if (!c || rep < -1 || rep >= c->n_rep) return fail(Fail{TIS_EINVAL, "tis_set_rng_epoch: bad arguments"});
    return guarded([&]() -> int {
        HIP_TRY(hipStreamSynchronize(c->s_main));
        for (int r = 0; r < c->n_rep; ++r) {
            if (rep >= 0 && r != rep) continue;
            RepState s = c->get_rs(r);
            s.rng_epoch = epoch;
            c->put_rs(r, s);
        }
        return TIS_OK;
    });
}

// Captures and instantiates the graph of opt.graph_steps steps starting from the
// current parity of the state buffers (launching nothing) unless it exists.
// Capture issues no work, so the parity is restored afterwards.
static void capture_graph(tis_ctx *c) {
    const int p0 = c->par;
    if (c->graph[p0]) return;
    hipGraph_t g;
    HIP_TRY(hipStreamBeginCapture(c->s_main, hipStreamCaptureModeThreadLocal));
    for (int s = 0; s < c->opt.graph_steps; ++s) c->launch_step();
    HIP_TRY(hipStreamEndCapture(c->s_main, &g));
    c->set_parity(p0);
    HIP_TRY(hipGraphInstantiate(&c->graph[p0], g, nullptr, nullptr, 0));
    HIP_TRY(hipGraphDestroy(g));
    c->graph_steps = c->opt.graph_steps;
}

extern "C" int tis_prepare_graph(tis_ctx *c, int graph_steps) {
    if (!c || graph_steps < 0) return fail(Fail{TIS_EINVAL, "tis_prepare_graph: bad arguments"});
    return guarded([&]() -> int {
        HIP_TRY(hipStreamSynchronize(c->s_main));
        if (graph_steps != c->graph_steps || !c->opt.use_graph || graph_steps == 0) c->drop_graphs();
        c->opt.graph_steps = graph_steps;
        if (graph_steps > 0 && c->opt.use_graph && !c->profiling) capture_graph(c);
        return TIS_OK;
    });
}

extern "C" int tis_set_persistent(tis_ctx *c, int on) {
    if (!c || (on != 0 && on != 1)) return fail(Fail{TIS_EINVAL, "tis_set_persistent: bad arguments"});
    if (on && !c->persist_ready)
        return fail(Fail{TIS_EINVAL, "tis_set_persistent: the persistent step is not available: " + c->persist_why});
    c->persist = on != 0;
    return TIS_OK;
}

extern "C" int tis_get_persistent_counts(tis_ctx *c, int *active, long long *launches, long long *steps,
                                         char *why, int why_len) {
    if (!c) return fail(Fail{TIS_EINVAL, "tis_get_persistent_counts: bad arguments"});
    if (active) *active = c->persist ? 1 : 0;
    if (launches) *launches = c->n_persist_launch;
    if (steps) *steps = c->n_persist_step;
    if (why && why_len > 0) {
        std::strncpy(why, c->persist_why.c_str(), (size_t)why_len - 1);
        why[why_len - 1] = 0;
    }
    return TIS_OK;
}

extern "C" int tis_trace_hb(tis_ctx *c, int on) {
    if (!c || (on != 0 && on != 1)) return fail(Fail{TIS_EINVAL, "tis_trace_hb: bad arguments"});
    return guarded([&]() -> int {
        HIP_TRY(hipStreamSynchronize(c->s_main));
        const size_t words = (size_t)kHbTraceSteps * c->n_rep * kMaxHbGroups * kHbTraceWords;
        if (on && !c->hb_trace_buf) c->hb_trace_buf = dalloc<unsigned long long>(c->owned, words);
        if (on) HIP_TRY(hipMemset(c->hb_trace_buf, 0, words * sizeof(unsigned long long)));
        c->D.hb_trace = on ? c->hb_trace_buf : nullptr;
        c->drop_graphs();   // captured graphs hold the old D
        return TIS_OK;
    });
}

extern "C" int tis_get_hb_trace(tis_ctx *c, unsigned long long *out, long long n) {
    const long long words = c ? (long long)kHbTraceSteps * c->n_rep * kMaxHbGroups * kHbTraceWords : 0;
    if (!c || !out || n < words) return fail(Fail{TIS_EINVAL, "tis_get_hb_trace: bad arguments"});
    if (!c->hb_trace_buf) return fail(Fail{TIS_EINVAL, "tis_get_hb_trace: the trace was never switched on"});
    return guarded([&]() -> int {
        HIP_TRY(hipStreamSynchronize(c->s_main));
        HIP_TRY(hipMemcpy(out, c->hb_trace_buf, words * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        return TIS_OK;
    });
}

extern "C" int tis_get_launch_counts(tis_ctx *c, long long *graph_launches, long long *graph_steps,
                                     long long *eager_steps) {
    if (!c) return fail(Fail{TIS_EINVAL, "tis_get_launch_counts: bad arguments"});
    if (graph_launches) *graph_launches = c->n_graph_launch;
    if (graph_steps) *graph_steps = c->n_graph_step;
    if (eager_steps) *eager_steps = c->n_eager_step;
    return TIS_OK;
}

// n steps: event-timed eager launches when profiling, else whole captured
// graphs of graph_steps and eager launches for the remainder.
static void run_steps(tis_ctx *c, long long n_steps) {
    long long done = 0;
    if (c->persist && !c->profiling && c->p_G >= c->hb_groups + 2) {
        // at most kPersistChunk steps per launch (barrier counters stay small)
        constexpr long long kPersistChunk = 20000;
        while (done < n_steps) {
            const long long m = std::min(kPersistChunk, n_steps - done);
            c->launch_persist(m);
            done += m;
        }
        if (c->persist_status() & PSTAT_TIMEOUT)
            throw Fail{TIS_EDEVICE, "persistent step: a barrier wait gave up (workgroups of a replica out of step)"};
        return;
    }
    if (c->profiling) {
        // bounded batches so that the event pool stays small
        while (done < n_steps) {
            const long long m = std::min<long long>(64, n_steps - done);
            for (long long q = 0; q < m; ++q) c->launch_step();
            HIP_TRY(hipStreamSynchronize(c->s_main));
            HIP_TRY(hipStreamSynchronize(c->s_hb));
            HIP_TRY(hipStreamSynchronize(c->s_mb));
            c->collect_times();
            done += m;
            c->n_eager_step += m;
        }
    }
    if (c->opt.use_graph && c->opt.graph_steps > 0 && n_steps - done >= c->opt.graph_steps) {
        capture_graph(c);
        while (n_steps - done >= c->graph_steps) {
            capture_graph(c);   // a graph for this parity (after an odd graph or remainder)
            HIP_TRY(hipGraphLaunch(c->graph[c->par], c->s_main));
            c->set_parity(c->par ^ (c->graph_steps & 1));
            done += c->graph_steps;
            c->n_graph_launch += 1;
            c->n_graph_step += c->graph_steps;
        }
    }
    c->n_eager_step += n_steps - done;
    for (; done < n_steps; ++done) c->launch_step();
}

extern "C" int tis_step(tis_ctx *c, long long n_steps) {
    if (!c || n_steps < 0) return fail(Fail{TIS_EINVAL, "tis_step: bad arguments"});
    return guarded([&]() -> int {
        const long long ex = c->opt.exchange_every, ob = c->obs_every;
        long long left = n_steps;
        while (left > 0) {
            long long chunk = left;
            if (ex > 0) chunk = std::min(chunk, ex - c->steps_done % ex);
            if (ob > 0) chunk = std::min(chunk, ob - c->steps_done % ob);
            run_steps(c, chunk);
            c->steps_done += chunk;
            left -= chunk;
            if (ex > 0 && c->steps_done % ex == 0) c->launch_exchange();
            if (ob > 0 && c->steps_done % ob == 0) c->launch_observe();
        }
        HIP_TRY(hipGetLastError());
        if (n_steps > 0) c->finish_pending_rebuild();
        HIP_TRY(hipStreamSynchronize(c->s_main));
        if (c->profiling) c->collect_times();
        c->check_flags();
        return TIS_OK;
    });
}

// ---- autotuner: launch shapes timed on a scratch context
static void apply_launch(tis_ctx *c, const tis_options &o) {
    check_launch_options(o);
    c->opt.pair_lanes = o.pair_lanes;
    c->opt.pair_block = o.pair_block;
    c->opt.list_waves = o.list_waves;
    c->opt.graph_steps = o.graph_steps;
    c->opt.hb_groups = o.hb_groups;
    c->hb_groups = o.hb_groups;
    c->D.hb_groups = o.hb_groups;
    c->D.pair_lanes = o.pair_lanes;
    c->pair_block = o.pair_block;
    c->list_waves = o.list_waves;
    c->drop_graphs();
}

extern "C" int tis_autotune(const tis_system *reps, int n_rep, tis_options *opt, int trial_steps, tis_tuning *report) {
    if (!reps || n_rep <= 0 || !opt || trial_steps <= 0) return fail(Fail{TIS_EINVAL, "tis_autotune: bad arguments"});
    tis_ctx *c = nullptr;
    const int rc = tis_create(reps, n_rep, opt, &c);
    if (rc != TIS_OK) return rc;
    const int out = guarded([&]() -> int {
        int trials = 0;
        auto measure = [&](const tis_options &t) {
            apply_launch(c, t);
            if (tis_step(c, trial_steps / 4 + 1) != TIS_OK) throw Fail{TIS_EDEVICE, tis_last_error()};
            double best = 1e300;
            for (int rep = 0; rep < 2; ++rep) {
                const auto t0 = std::chrono::steady_clock::now();
                if (tis_step(c, trial_steps) != TIS_OK) throw Fail{TIS_EDEVICE, tis_last_error()};
                const std::chrono::duration<double, std::milli> dt = std::chrono::steady_clock::now() - t0;
                best = std::min(best, dt.count() / trial_steps);
            }
            ++trials;
            return best;
        };
        tis_options best = *opt;
        const double start = measure(best);
        double best_ms = start;
        // coordinate descent; a candidate must win by 1 % to replace the incumbent
        auto sweep = [&](int tis_options::*field, std::initializer_list<int> cands) {
            for (int v : cands) {
                if (v == best.*field) continue;
                tis_options t = best;
                t.*field = v;
                const double ms = measure(t);
                if (ms < 0.99 * best_ms) { best = t; best_ms = ms; }
            }
        };
        if (best.use_graph) sweep(&tis_options::graph_steps, {25, 50, 100, 200});
        sweep(&tis_options::pair_lanes, {2, 4, 8, 16});
        sweep(&tis_options::pair_block, {64, 128, 256});
        sweep(&tis_options::list_waves, {1024, 2048, 4096, 8192, 16384});
        if (best.rng == TIS_RNG_PHILOX) sweep(&tis_options::hb_groups, {1, 2, 4, 8});
        opt->graph_steps = best.graph_steps;
        opt->pair_lanes = best.pair_lanes;
        opt->pair_block = best.pair_block;
        opt->list_waves = best.list_waves;
        opt->hb_groups = best.hb_groups;
        if (report) {
            report->ms_per_step_start = start;
            report->ms_per_step_best = best_ms;
            report->trials = trials;
        }
        return TIS_OK;
    });
    tis_destroy(c);
    return out;
}

extern "C" int tis_set_force_field(tis_ctx *c, const double *hb_E, const double *st3_E, const double *st2_E,
                                   const double *ion_R) {
    if (!c) return fail(Fail{TIS_EINVAL, "tis_set_force_field: bad arguments"});
    if ((hb_E || st3_E || st2_E) && !c->all_same)
        return fail(Fail{TIS_EINVAL, "tis_set_force_field: per-term parameters need replicas of one system"});
    return guarded([&]() -> int {
        HIP_TRY(hipStreamSynchronize(c->s_main));
        Dev &D = c->D;
        // per-term sets: replica 0's values uploaded once and copied to every
        // replica's region on the device (the replicas describe one system)
        auto set = [&](double *dst, int n, int stride, int at, const double *src) {
            if (!src || n <= 0) return;
            HIP_TRY(hipMemcpyAsync(c->fm_small, src, n * sizeof(double), hipMemcpyHostToDevice, c->s_main));
            const long long tot = (long long)n * c->n_rep;
            k_set_param<<<(int)((tot + kBlock - 1) / kBlock), kBlock, 0, c->s_main>>>(dst, n, c->n_rep, stride, at,
                                                                                     c->fm_small);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(c->s_main));
        };
        const RepParams &P0 = c->rp[0];
        set(D.hb_E, P0.nhb, 1, 0, hb_E);
        set(D.st3_E, P0.nst3, 1, 0, st3_E);
        set((double *)D.st2_rE, P0.nst2, 2, 1, st2_E);
        if (ion_R)
            for (int r = 0; r < c->n_rep; ++r) {
                for (int t = 0; t < 3; ++t) c->type_R[r][6 + t] = ion_R[t];
                type_tables(c->rp[r], c->type_R[r].data(), c->type_eps[r].data(), c->type_q[r].data(),
                            c->type_mass[r].data(), c->rp[r].lB);
            }
        if (ion_R) HIP_TRY(hipMemcpy(D.rp, c->rp.data(), c->n_rep * sizeof(RepParams), hipMemcpyHostToDevice));
        return TIS_OK;
    });
}

extern "C" int tis_force_match(tis_ctx *c, const double *target, double *loss, double *forces, double *grad_pos,
                               double *grad_hb_E, double *grad_st3_E, double *grad_st2_E, double *grad_ion_R) {
    if (!c || !target || !loss) return fail(Fail{TIS_EINVAL, "tis_force_match: bad arguments"});
    if ((grad_hb_E || grad_st3_E || grad_st2_E) && !c->all_same)
        return fail(Fail{TIS_EINVAL, "tis_force_match: per-term gradients need replicas of one system"});
    return guarded([&]() -> int {
        Dev &D = c->D;
        const int N = D.n_tot;
        hipStream_t st = c->s_main;
        HIP_TRY(hipStreamSynchronize(st));
        HIP_TRY(hipMemcpyAsync(c->fm_t3, target, 3 * (size_t)N * sizeof(double), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemsetAsync(c->fm_loss, 0, sizeof(double), st));
        HIP_TRY(hipMemsetAsync(c->fm_ion, 0, 3 * sizeof(double), st));
        HIP_TRY(hipMemsetAsync(c->fm_ghb, 0, std::max<size_t>(1, c->h_hbE.size()) * sizeof(double), st));
        c->launch_force_match(st, c->fm_t3, c->fm_forces, nullptr, c->fm_loss, c->fm_ion, c->fm_ghb, c->fm_gst3,
                              c->fm_gst2, 0);
        HIP_TRY(hipStreamSynchronize(st));
        c->check_flags();
        const double M = 3.0 * N, sc = 2.0 / M;
        double l = 0;
        HIP_TRY(hipMemcpy(&l, c->fm_loss, sizeof(double), hipMemcpyDeviceToHost));
        *loss = l / M;
        // device results packed on the context's stream, then copied on that stream
        // (the stream does not synchronise with the null stream)
        auto fetch = [&](double *out, const double *dev, size_t n) {
            HIP_TRY(hipMemcpyAsync(out, dev, n * sizeof(double), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
        };
        if (forces) {
            k_pack3<<<c->grid(N), kBlock, 0, st>>>(c->fm_forces, N, 1.0, c->fm_out3);
            fetch(forces, c->fm_out3, 3 * (size_t)N);
        }
        if (grad_pos) {
            k_pack3<<<c->grid(N), kBlock, 0, st>>>(D.jf, N, sc, c->fm_out3);
            fetch(grad_pos, c->fm_out3, 3 * (size_t)N);
        }
        // per-term gradients: summed over the replicas (one system) on the device
        auto summed = [&](const double *dev, int n_local, double *out) {
            if (n_local <= 0) return;
            k_sum_replicas<<<(n_local + kBlock - 1) / kBlock, kBlock, 0, st>>>(dev, n_local, c->n_rep, sc, c->fm_small);
            fetch(out, c->fm_small, (size_t)n_local);
        };
        if (grad_hb_E) summed(c->fm_ghb, c->rp[0].nhb, grad_hb_E);
        if (grad_st3_E) summed(c->fm_gst3, c->rp[0].nst3, grad_st3_E);
        if (grad_st2_E) summed(c->fm_gst2, c->rp[0].nst2, grad_st2_E);
        if (grad_ion_R) {
            double g3[3];
            HIP_TRY(hipMemcpy(g3, c->fm_ion, 3 * sizeof(double), hipMemcpyDeviceToHost));
            for (int t = 0; t < 3; ++t) grad_ion_R[t] = sc * g3[t];
        }
        return TIS_OK;
    });
}

extern "C" int tis_force_match_noalloc(tis_ctx *c, const double *target, double *loss, double *forces,
                                       double *grad_pos, double *grad_hb_E, double *grad_st3_E, double *grad_st2_E,
                                       double *grad_ion_R) {
    if (!c || !target || !loss) return fail(Fail{TIS_EINVAL, "tis_force_match_noalloc: bad arguments"});
    if ((grad_hb_E || grad_st3_E || grad_st2_E) && !c->all_same)
        return fail(Fail{TIS_EINVAL, "tis_force_match_noalloc: per-term gradients need replicas of one system"});
    // sums at replica-local term indices: regions as large as the largest
    // replica's term counts (per-term outputs need one system, where all agree)
    const RepParams &P0 = c->rp[0];
    int mhb = 0, mst3 = 0, mst2 = 0;
    for (const RepParams &q : c->rp) {
        mhb = std::max(mhb, q.nhb);
        mst3 = std::max(mst3, q.nst3);
        mst2 = std::max(mst2, q.nst2);
    }
    const long long need = 4LL + mhb + mst3 + mst2;
    if (need > kScratchDoubles)
        return fail(Fail{TIS_EINVAL, "tis_force_match_noalloc: " + std::to_string(need) +
                                     " working sums (4 + HB candidates + stacks) exceed the " +
                                     std::to_string(kScratchDoubles) + " of the device scratch"});
    return guarded([&]() -> int {
        Dev &D = c->D;
        const int N = D.n_tot;
        hipStream_t st = c->s_main;
        double *w = c->scratch_dev;
        // every call starts from zero sums (the array keeps the last call's values)
        k_fill_zero<<<(int)((need + kBlock - 1) / kBlock), kBlock, 0, st>>>(w, need);
        double *whb = w + 4, *wst3 = whb + mhb, *wst2 = wst3 + mst3;
        c->launch_force_match(st, target, nullptr, forces, w, w + 1, whb, wst3, wst2, 1);
        const double M = 3.0 * N, sc = 2.0 / M;
        if (grad_pos) k_pack3<<<c->grid(N), kBlock, 0, st>>>(D.jf, N, sc, grad_pos);
        const int nfin = std::max({3, P0.nhb, P0.nst3, P0.nst2});
        k_fm_finish<<<(nfin + kBlock - 1) / kBlock, kBlock, 0, st>>>(w, whb, wst3, wst2, P0.nhb, P0.nst3, P0.nst2, sc,
                                                                      1.0 / M, loss, grad_ion_R, grad_hb_E, grad_st3_E,
                                                                      grad_st2_E);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(st));
        c->check_flags();
        return TIS_OK;
    });
}

extern "C" int tis_exchange_now(tis_ctx *c) {
    if (!c) return fail(Fail{TIS_EINVAL, "tis_exchange_now: bad arguments"});
    return guarded([&]() -> int {
        HIP_TRY(hipStreamSynchronize(c->s_main));
        c->launch_exchange();
        HIP_TRY(hipStreamSynchronize(c->s_main));
        c->check_flags();
        return TIS_OK;
    });
}

extern "C" int tis_get_exchange_stats(tis_ctx *c, long long *attempts, long long *accepts, double *last_d,
                                      int n_pairs) {
    if (!c || n_pairs < 0 || n_pairs > std::max(0, c->n_rep - 1))
        return fail(Fail{TIS_EINVAL, "tis_get_exchange_stats: bad arguments"});
    return guarded([&]() -> int {
        HIP_TRY(hipStreamSynchronize(c->s_main));
        std::vector<long long> st(2 * (size_t)c->n_rep);
        std::vector<double> d(c->n_rep);
        HIP_TRY(hipMemcpy(st.data(), c->D.remd_stats, st.size() * sizeof(long long), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(d.data(), c->D.remd_d, d.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int q = 0; q < n_pairs; ++q) {
            if (attempts) attempts[q] = st[2 * q];
            if (accepts) accepts[q] = st[2 * q + 1];
            if (last_d) last_d[q] = d[q];
        }
        return TIS_OK;
    });
}

extern "C" int tis_get_walkers(tis_ctx *c, int *walker, int n_rep) {
    if (!c || !walker || n_rep != c->n_rep) return fail(Fail{TIS_EINVAL, "tis_get_walkers: bad arguments"});
    return guarded([&]() -> int {
        HIP_TRY(hipStreamSynchronize(c->s_main));
        HIP_TRY(hipMemcpy(walker, c->D.walker, n_rep * sizeof(int), hipMemcpyDeviceToHost));
        return TIS_OK;
    });
}

extern "C" int tis_observe(tis_ctx *c, long long every, double cut) {
    if (!c || every < 0 || !(cut > 0)) return fail(Fail{TIS_EINVAL, "tis_observe: bad arguments"});
    const double lim = c->opt.r_sr - c->opt.skin - 1.0;
    if (cut > lim)
        return fail(Fail{TIS_EINVAL, "tis_observe: contact cut above r_sr - skin - 1 = " + std::to_string(lim) +
                                     " A (pairs beyond it may be missing from the short-range lists)"});
    return guarded([&]() -> int {
        c->obs_every = every;
        c->obs_cut = cut;
        c->reset_observables();
        HIP_TRY(hipStreamSynchronize(c->s_main));
        return TIS_OK;
    });
}

extern "C" int tis_observe_now(tis_ctx *c) {
    if (!c) return fail(Fail{TIS_EINVAL, "tis_observe_now: bad arguments"});
    return guarded([&]() -> int {
        c->launch_observe();
        HIP_TRY(hipStreamSynchronize(c->s_main));
        c->check_flags();
        return TIS_OK;
    });
}

extern "C" int tis_get_observables(tis_ctx *c, int rep, tis_observables *o, double *per_bead, int n_beads,
                                   double *hb_kept, int n_hb, double *st3_mean, int n_st3) {
    if (!c || rep < 0 || rep >= c->n_rep || !o) return fail(Fail{TIS_EINVAL, "tis_get_observables: bad arguments"});
    const RepParams &P = c->rp[rep];
    if ((per_bead && n_beads != P.nbeads) || (hb_kept && n_hb != P.nhb) || (st3_mean && n_st3 != P.nst3))
        return fail(Fail{TIS_EINVAL, "tis_get_observables: array sizes differ from the replica's"});
    return guarded([&]() -> int {
        HIP_TRY(hipStreamSynchronize(c->s_main));
        double a[kObsRep];
        HIP_TRY(hipMemcpy(a, c->D.obs_rep + (size_t)kObsRep * rep, sizeof a, hipMemcpyDeviceToHost));
        const double ns = a[0], w = ns > 0 ? 1.0 / ns : 0.0;
        std::memset(o, 0, sizeof *o);
        o->n_samples = (long long)ns;
        o->cut = c->obs_cut;
        o->rg = a[1] * w; o->rg2 = a[2] * w; o->asphericity = a[3] * w; o->shape = a[4] * w;
        o->e2_hb = a[5] * w; o->e3_hb = a[6] * w; o->e2_stack = a[7] * w; o->e3_stack = a[8] * w;
        o->e_wca = a[9] * w; o->e_coul_sr = a[10] * w; o->e_coul_lr = a[11] * w;
        o->e_bond = a[12] * w; o->e_angle = a[13] * w;
        o->bound_mg = a[14] * w; o->bound_cl = a[15] * w; o->bound_k = a[16] * w;
        if (per_bead) {
            std::vector<double4> v(P.nbeads);
            HIP_TRY(hipMemcpy(v.data(), c->D.obs_p + P.p0, P.nbeads * sizeof(double4), hipMemcpyDeviceToHost));
            for (int k = 0; k < P.nbeads; ++k) {
                per_bead[4 * k + 0] = v[k].x * w; per_bead[4 * k + 1] = v[k].y * w;
                per_bead[4 * k + 2] = v[k].z * w; per_bead[4 * k + 3] = v[k].w * w;
            }
        }
        if (hb_kept && P.nhb > 0) {
            HIP_TRY(hipMemcpy(hb_kept, c->D.obs_hb + P.hb0, P.nhb * sizeof(double), hipMemcpyDeviceToHost));
            for (int k = 0; k < P.nhb; ++k) hb_kept[k] *= w;
        }
        if (st3_mean && P.nst3 > 0) {
            HIP_TRY(hipMemcpy(st3_mean, c->D.obs_st3 + P.st30, P.nst3 * sizeof(double), hipMemcpyDeviceToHost));
            for (int k = 0; k < P.nst3; ++k) st3_mean[k] *= w;
        }
        return TIS_OK;
    });
}

extern "C" int tis_get_energies(tis_ctx *c, int rep, tis_energies *e) {
    if (!c || rep < 0 || rep >= c->n_rep || !e) return fail(Fail{TIS_EINVAL, "tis_get_energies: bad arguments"});
    return guarded([&]() -> int {
        HIP_TRY(hipStreamSynchronize(c->s_main));
        RepState s = c->get_rs(rep);
        const RepParams &P = c->rp[rep];
        std::memset(e, 0, sizeof *e);
        e->e2_hb = s.e[EI_E2_HB]; e->e3_hb = s.e[EI_E3_HB];
        e->e2_stack = s.e[EI_E2_ST]; e->e3_stack = s.e[EI_E3_ST];
        e->e_wca = s.e[EI_WCA]; e->e_coul_sr = s.e[EI_COUL_SR]; e->e_coul_lr = s.e[EI_COUL_LR];
        e->e_bond = s.e[EI_BOND]; e->e_angle = s.e[EI_ANGLE];
        e->klok = s.klok; e->n_rebuilds = s.n_rebuilds; e->progress = s.progress;
        e->n_hb_active = s.n_act;
        e->hb_rounds = s.hb_iters;
        // Rg of the beads (radius_of_gyration, fast.cpp:2844-2876)
        std::vector<double4> x(P.nbeads);
        HIP_TRY(hipMemcpy(x.data(), c->D.pos + P.p0, P.nbeads * sizeof(double4), hipMemcpyDeviceToHost));
        double cm[3] = {0, 0, 0};
        for (const double4 &q : x) { cm[0] += q.x; cm[1] += q.y; cm[2] += q.z; }
        for (double &v : cm) v /= P.nbeads;
        double acc = 0;
        for (const double4 &q : x) {
            const double a = q.x - cm[0], b = q.y - cm[1], d = q.z - cm[2];
            acc += a * a + b * b + d * d;
        }
        e->rg = std::sqrt(acc / P.nbeads);
        // pair counts from the list lengths
        std::vector<int> nsr(P.n), nlr(P.n);
        HIP_TRY(hipMemcpy(nsr.data(), c->D.nb_nsr + P.p0, P.n * sizeof(int), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(nlr.data(), c->D.nb_nlr + P.p0, P.n * sizeof(int), hipMemcpyDeviceToHost));
        long long a = 0, b = 0;
        for (int i = 0; i < P.n; ++i) { a += nsr[i]; b += nlr[i]; }
        e->n_sr_pairs = a / 2;
        e->n_lr_pairs = b / 2;
        return TIS_OK;
    });
}

extern "C" int tis_get_stack_energies(tis_ctx *c, int rep, double *out, int n) {
    if (!c || rep < 0 || rep >= c->n_rep || !out) return fail(Fail{TIS_EINVAL, "tis_get_stack_energies: bad arguments"});
    return guarded([&]() -> int {
        const RepParams &P = c->rp[rep];
        if (n < P.nst3) throw Fail{TIS_EINVAL, "tis_get_stack_energies: buffer too small"};
        HIP_TRY(hipStreamSynchronize(c->s_main));
        if (P.nst3) HIP_TRY(hipMemcpy(out, c->D.st_energy + P.st30, P.nst3 * sizeof(double), hipMemcpyDeviceToHost));
        return TIS_OK;
    });
}

// Forces of the selected terms at the current positions of replica rep.  The
// other replicas are left untouched (their terms run but write nothing that
// the step keeps: multi-body forces are cleared afterwards, and the SR/LR
// buffers are recomputed by the next step).
extern "C" int tis_eval_forces(tis_ctx *c, int rep, unsigned mask, double *f_out, tis_energies *e) {
    if (!c || rep < 0 || rep >= c->n_rep) return fail(Fail{TIS_EINVAL, "tis_eval_forces: bad arguments"});
    return guarded([&]() -> int {
        Dev &D = c->D;
        const RepParams &P = c->rp[rep];
        HIP_TRY(hipStreamSynchronize(c->s_main));
        // clear energies and forces
        std::vector<RepState> all(c->n_rep);
        HIP_TRY(hipMemcpy(all.data(), D.rs, c->n_rep * sizeof(RepState), hipMemcpyDeviceToHost));
        for (auto &s : all) std::memset(s.e, 0, sizeof s.e);
        HIP_TRY(hipMemcpy(D.rs, all.data(), c->n_rep * sizeof(RepState), hipMemcpyHostToDevice));
        HIP_TRY(hipMemsetAsync(D.f_mb, 0, D.n_tot * sizeof(double4), c->s_main));
        if (mask & TIS_TERM_HB) {
            c->launch_hb(c->s_main, 0);
        }
        const unsigned m = mask & (TIS_TERM_STACK | TIS_TERM_BACKBONE | TIS_TERM_SR | TIS_TERM_LR);
        const int nmb = c->n_rep + c->nb_bonded, npr = c->pair_grid();
        if (D.precision == TIS_PREC_FP64) {
            k_force<double, 0><<<nmb, kBlock, 0, c->s_main>>>(D, c->nb_bonded, m, kEvalFull);
            k_force<double, 1><<<npr, c->pair_block, 0, c->s_main>>>(D, c->nb_bonded, m, kEvalFull);
        } else {
            k_force<float, 0><<<nmb, kBlock, 0, c->s_main>>>(D, c->nb_bonded, m, kEvalFull);
            k_force<float, 1><<<npr, c->pair_block, 0, c->s_main>>>(D, c->nb_bonded, m, kEvalFull);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(c->s_main));
        std::vector<double4> a(P.n), b(P.n), d(P.n);
        HIP_TRY(hipMemcpy(a.data(), D.f_sr + P.p0, P.n * sizeof(double4), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(b.data(), D.f_lr + P.p0, P.n * sizeof(double4), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(d.data(), D.f_mb + P.p0, P.n * sizeof(double4), hipMemcpyDeviceToHost));
        const double ws = (mask & TIS_TERM_SR) ? 1.0 : 0.0, wl = (mask & TIS_TERM_LR) ? 1.0 : 0.0;
        if (f_out)
            for (int i = 0; i < P.n; ++i) {
                f_out[3 * i] = ws * a[i].x + wl * b[i].x + d[i].x;
                f_out[3 * i + 1] = ws * a[i].y + wl * b[i].y + d[i].y;
                f_out[3 * i + 2] = ws * a[i].z + wl * b[i].z + d[i].z;
            }
        HIP_TRY(hipMemsetAsync(D.f_mb, 0, D.n_tot * sizeof(double4), c->s_main));
        HIP_TRY(hipStreamSynchronize(c->s_main));
        c->check_flags();
        if (e) {
            int rc = tis_get_energies(c, rep, e);
            if (rc) return rc;
        }
        return TIS_OK;
    });
}

extern "C" int tis_get_pairs(tis_ctx *c, int rep, int which, int *ij_out, long long cap, long long *n_out) {
    if (!c || rep < 0 || rep >= c->n_rep || (which != 0 && which != 1) || !n_out)
        return fail(Fail{TIS_EINVAL, "tis_get_pairs: bad arguments"});
    return guarded([&]() -> int {
        const RepParams &P = c->rp[rep];
        HIP_TRY(hipStreamSynchronize(c->s_main));
        std::vector<int> st(P.n), ns(P.n), nl(P.n);
        HIP_TRY(hipMemcpy(st.data(), c->D.nb_start + P.p0, P.n * sizeof(int), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(ns.data(), c->D.nb_nsr + P.p0, P.n * sizeof(int), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(nl.data(), c->D.nb_nlr + P.p0, P.n * sizeof(int), hipMemcpyDeviceToHost));
        std::vector<int> pool((size_t)P.pool_cap), sorted(P.n);
        HIP_TRY(hipMemcpy(pool.data(), c->D.pool + P.pool0, pool.size() * sizeof(int), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(sorted.data(), c->D.sorted + P.p0, P.n * sizeof(int), hipMemcpyDeviceToHost));
        std::vector<std::pair<int, int>> pr;
        for (int i = 0; i < P.n; ++i) {
            const int b = st[i] + (which ? P.cap_sr : 0), n = which ? nl[i] : ns[i];
            for (int q = 0; q < n; ++q) {
                const int e = pool[b + q];   // cell-order index of the partner
                const int j = sorted[(e < 0 ? ~e : e) - P.p0] - P.p0;
                if (i < j) pr.push_back({i, j});
            }
        }
        std::sort(pr.begin(), pr.end());
        *n_out = (long long)pr.size();
        if (ij_out)
            for (long long q = 0; q < std::min<long long>(cap, (long long)pr.size()); ++q) {
                ij_out[2 * q] = pr[q].first;
                ij_out[2 * q + 1] = pr[q].second;
            }
        return TIS_OK;
    });
}

extern "C" int tis_get_hb_status(tis_ctx *c, int rep, int *out, int n) {
    if (!c || rep < 0 || rep >= c->n_rep || !out) return fail(Fail{TIS_EINVAL, "tis_get_hb_status: bad arguments"});
    return guarded([&]() -> int {
        const RepParams &P = c->rp[rep];
        if (n < P.nhb) throw Fail{TIS_EINVAL, "tis_get_hb_status: buffer too small"};
        HIP_TRY(hipStreamSynchronize(c->s_main));
        std::vector<signed char> s(P.nhb);
        if (P.nhb) HIP_TRY(hipMemcpy(s.data(), c->D.hb_status + P.hb0, P.nhb, hipMemcpyDeviceToHost));
        for (int k = 0; k < P.nhb; ++k) out[k] = s[k];
        return TIS_OK;
    });
}

extern "C" int tis_get_hb_outcome(tis_ctx *c, int rep, int *out, int n) {
    if (!c || rep < 0 || rep >= c->n_rep || !out) return fail(Fail{TIS_EINVAL, "tis_get_hb_outcome: bad arguments"});
    return guarded([&]() -> int {
        const RepParams &P = c->rp[rep];
        if (n < P.nhb) throw Fail{TIS_EINVAL, "tis_get_hb_outcome: buffer too small"};
        HIP_TRY(hipStreamSynchronize(c->s_main));
        HIP_TRY(hipStreamSynchronize(c->s_hb));
        std::vector<signed char> v(P.nhb);
        std::vector<signed char> st(P.nhb);
        if (P.nhb) HIP_TRY(hipMemcpy(v.data(), c->D.hb_out + P.hb0, P.nhb, hipMemcpyDeviceToHost));
        if (P.nhb) HIP_TRY(hipMemcpy(st.data(), c->D.hb_status + P.hb0, P.nhb, hipMemcpyDeviceToHost));
        // listed by the latest rebuild but not by the one the last resolution read
        // (tables built beside it): outside its gate by the rule of launch_step
        for (int k = 0; k < P.nhb; ++k) out[k] = st[k] < 0 ? -1 : (v[k] < 0 ? 0 : v[k]);
        return TIS_OK;
    });
}

extern "C" int tis_set_uniform_stream(tis_ctx *c, int rep, const int32_t *raw, long long n) {
    if (!c || rep < 0 || rep >= c->n_rep || n < 0 || (n && !raw))
        return fail(Fail{TIS_EINVAL, "tis_set_uniform_stream: bad arguments"});
    return guarded([&]() -> int {
        if (c->opt.rng != TIS_RNG_STREAM) throw Fail{TIS_ESTATE, "context was created without TIS_RNG_STREAM"};
        HIP_TRY(hipStreamSynchronize(c->s_main));
        // regions are re-laid out: replica rep gets a fresh allocation of n values
        if (n > c->stream_cap[rep]) {
            std::vector<long long> base(c->n_rep), len(c->n_rep);
            HIP_TRY(hipMemcpy(base.data(), c->D.stream_base, c->n_rep * sizeof(long long), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(len.data(), c->D.stream_len, c->n_rep * sizeof(long long), hipMemcpyDeviceToHost));
            long long total = c->stream_total + n;
            int32_t *ns = nullptr;
            HIP_TRY(hipMalloc(&ns, std::max<long long>(total, 1) * sizeof(int32_t)));
            if (c->D.stream && c->stream_total)
                HIP_TRY(hipMemcpy(ns, c->D.stream, c->stream_total * sizeof(int32_t), hipMemcpyDeviceToDevice));
            if (c->D.stream) {
                auto it = std::find(c->owned.begin(), c->owned.end(), (void *)c->D.stream);
                if (it != c->owned.end()) c->owned.erase(it);
                HIP_TRY(hipFree(c->D.stream));
            }
            c->owned.push_back(ns);
            c->D.stream = ns;
            base[rep] = c->stream_total;
            c->stream_total = total;
            c->stream_cap[rep] = n;
            HIP_TRY(hipMemcpy(c->D.stream_base, base.data(), c->n_rep * sizeof(long long), hipMemcpyHostToDevice));
            c->drop_graphs();
        }
        long long base_r;
        HIP_TRY(hipMemcpy(&base_r, c->D.stream_base + rep, sizeof base_r, hipMemcpyDeviceToHost));
        if (n) HIP_TRY(hipMemcpy(c->D.stream + base_r, raw, n * sizeof(int32_t), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(c->D.stream_len + rep, &n, sizeof n, hipMemcpyHostToDevice));
        RepState s = c->get_rs(rep);
        s.stream_cursor = 0;
        s.overflow &= ~4;
        c->put_rs(rep, s);
        return TIS_OK;
    });
}

extern "C" int tis_get_stream_cursor(tis_ctx *c, int rep, long long *consumed) {
    if (!c || rep < 0 || rep >= c->n_rep || !consumed) return fail(Fail{TIS_EINVAL, "tis_get_stream_cursor: bad arguments"});
    return guarded([&]() -> int {
        HIP_TRY(hipStreamSynchronize(c->s_main));
        *consumed = c->get_rs(rep).stream_cursor;
        return TIS_OK;
    });
}

extern "C" int tis_kernel_times(tis_ctx *c, char *buf, int buflen) {
    if (!c || !buf || buflen <= 0) return fail(Fail{TIS_EINVAL, "tis_kernel_times: bad arguments"});
    std::string s;
    for (size_t k = 0; k < c->knames.size(); ++k) {
        char line[160];
        std::snprintf(line, sizeof line, "%s %lld %.6f\n", c->knames[k].c_str(), c->kcount[k],
                      c->kcount[k] ? c->ksum_ms[k] / c->kcount[k] : 0.0);
        s += line;
    }
    std::snprintf(buf, buflen, "%s", s.c_str());
    return TIS_OK;
}

extern "C" int tis_set_profiling(tis_ctx *c, int on) {
    if (!c) return fail(Fail{TIS_EINVAL, "tis_set_profiling: null context"});
    return guarded([&]() -> int {
        HIP_TRY(hipStreamSynchronize(c->s_main));
        c->profiling = on != 0;
        c->knames.clear();
        c->ksum_ms.clear();
        c->kcount.clear();
        c->pending.clear();
        c->event_next = 0;
        return TIS_OK;
    });
}

// Philox4x32-10 on the host, for known-answer tests of the device generator.
extern "C" void tis_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    Philox4 o = philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = o.v[i];
}
