// Persistent step: a batch's steps in one cooperative launch, each replica
// advancing on workgroups of its own, with no synchronisation between replicas.
//
// The step order is the reference's loop (fast.cpp:6380-6386: forces,
// noise, integration, check_shifts; the list rebuild of populate_lists at the
// start of a step whose trigger fired), as launch_step runs it with separate
// kernels; the phases are the same device bodies.  What changes is the
// synchronisation: launch_step separates its phases by kernel boundaries, which
// join every replica at every phase (a step lasts as long as the slowest
// replica's HB resolution) and cost a queue dependency each; here the
// workgroups of one replica separate its phases with barriers of their own, so
// each replica runs at its own pace.
//
// Placement.  One workgroup of kPBlock threads per CU (the resolution needs
// kHbLds of LDS), 8 XCDs of kCuPerXcd CUs; workgroup w runs on XCD w % 8
// (checked at launch, below).  Replica r = 8 j + x takes the workgroups of XCD x
// with (w / 8) / G == j: its G workgroups share one L2.  A store reaches that
// L2 before its workgroup arrives at a barrier (s_waitcnt), and every load of
// data another workgroup writes during the launch bypasses the CU's L1
// (LdCoh), so a barrier among the replica's workgroups orders their accesses
// without cache write-back or invalidation.  At launch each workgroup compares
// its XCC_ID with that of workgroup w % 8 and the launch does nothing if any
// differs (PStat bit 2); the host then keeps to launch_step.
//
// Roles of member m of G: m < H run the HB energies and the valence
// resolution (group m of H); the others the HB-table rebuild (when the lists
// are rebuilt), the multi-body slices, the random forces and the pair gather.
// All G: list rebuild (member G - 1 sorts); the non-HB members integrate and
// member H closes the step (commit_replica).
#pragma once
#include "kernels.h"

namespace tis {

constexpr int kPBlock = 1024;
constexpr int kXcds = 8;
constexpr int kPSpin = 1 << 20;    // barrier polls before a wait gives up (far longer than any phase)
constexpr int kPCtr = 32;          // ints per barrier counter (one 128-byte line each)
enum { PSTAT_TIMEOUT = 1, PSTAT_PLACEMENT = 2 };

// The two buffer-set descriptions of a launch (see k_persist) live in constant
// memory, kPSlots context slots of two each: constant loads are invariant for
// the compiler, which may repeat them instead of keeping the many pointers of
// the description live across the step.
constexpr int kPSlots = 16;
__constant__ Dev c_pdev[2 * kPSlots];

struct PArgs {
    int slot;        // this context's slot of c_pdev
    int *bar;        // barrier counters: 3 per replica (all, HB, others), then the launch-wide one
    int *status;     // PSTAT_* bits
    int *xcc;        // XCC_ID of each workgroup
    int n_steps;
    int G, q, H;     // workgroups per replica, replicas per XCD, HB workgroups per replica
};

TIS_DEV int ld_agent_i(const int *p) { return __hip_atomic_load(const_cast<int *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Barrier of the workgroups sharing counter c: every thread's stores and atomics
// completed (s_waitcnt), then thread 0 arrives and waits until the counter
// reaches target.  A wait gives up after kPSpin polls (PSTAT_TIMEOUT), and every
// wait ends once any status bit is set; false then.
TIS_DEV bool pbarrier(int *c, int target, int *status, int *sh_ok) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(c, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        for (int it = 0; ld_agent_i(c) < target; ++it) {
            if ((it & 63) == 63 && ld_agent_i(status) != 0) { ok = 0; break; }
            if (it > kPSpin) { atomicOr(status, PSTAT_TIMEOUT); ok = 0; break; }
            __builtin_amdgcn_s_sleep(1);
        }
        *sh_ok = ok;
    }
    __syncthreads();
    return *sh_ok != 0;
}

// Barrier counters of one replica, with the number of barriers passed so far.
struct PBars {
    int *all, *hb, *oth;
    int n_all, n_hb, n_oth;
};

// One step of replica r for member m.  The two roles run in separate step
// loops (k_persist), so each role's code and registers are its own.
template <class R, bool HB>
TIS_DEV bool persist_step(const Dev &D, const PArgs &A, int r, int m, unsigned char *lds, PBars &B, int *sh_ok) {
    RepState &S = D.rs[r];
    const RepParams &P = D.rp[r];
    const int G = A.G, H = A.H, tid = threadIdx.x, wv = tid >> 6;
    auto bar_all = [&] { return pbarrier(B.all, ++B.n_all * G, A.status, sh_ok); };
    const bool rebuild = LdCoh::ld(&S.need_rebuild) != 0;
    // ---- list rebuild (the trigger fired at the end of the last step)
    if (rebuild) {
        if (!HB && m == G - 1) bin_replica<LdCoh>(D, r, (int *)lds, (int *)lds + 2 * kMaxCells,
                                                   (unsigned short *)((int *)lds + 2 * kMaxCells + 32),
                                                   (int)((kHbLds - (2 * kMaxCells + 32) * sizeof(int)) / 2));
        if (!bar_all()) return false;
        list_replica<LdCoh, 1>(D, r, m * (kPBlock / 64) + wv, G * (kPBlock / 64));
        if (!bar_all()) return false;
    }
    if (HB) {
        // ---- HB energies, then the valence resolution (group m of H)
        const int n_act = (LdCoh::ld(&S.overflow) & 2) ? 0 : LdCoh::ld(&S.n_act);
        const int tab = LdCoh::ld(&S.tab);
        for (int a = m * kPBlock + tid; a < n_act; a += H * kPBlock) hbenergy_one<R, LdCoh>(D, r, a, n_act, tab);
        if (!pbarrier(B.hb, ++B.n_hb * H, A.status, sh_ok)) return false;
        hbres_group<R, LdCoh>(D, 1, r, m, H, lds);
        __syncthreads();
    } else {
        const int mo = m - H, Go = G - H;
        if (rebuild) {   // HB tables of the new lists, read from the next step on
            const int nch = (P.nhb + kStatusBlock - 1) / kStatusBlock;
            for (int c = mo; c < nch; c += Go) {
                hbstatus_chunk<LdCoh>(D, 1, r, c, (int *)lds);
                __syncthreads();
            }
            if (!pbarrier(B.oth, ++B.n_oth * Go, A.status, sh_ok)) return false;
            if (mo == Go - 1) {
                int *seg = (int *)lds, *aoff = seg + kAIncMax, *coff = aoff + kASiteMax + 1, *scr = coff + kMaxHbChunks + 1;
                hbsites_replica<LdCoh>(D, 1, r, HbSitesLds{scr + 32, scr, coff, aoff, seg});
                __syncthreads();
            }
        }
        // multi-body slices: bead forces in LDS, off flags and block sums after them
        {
            const int ns = D.st_slices;
            double *lf = (double *)lds;
            double *red = (double *)(lds + kHbLds) - kPBlock / 64;
            unsigned char *off = (unsigned char *)red - 4096;
            for (int sl = mo; sl < ns; sl += Go) {
                multibody_slice<R, LdCoh>(D, r, sl, ns, lf, off, red);
                __syncthreads();
            }
        }
        // random forces of the step
        {
            const long long klok = LdCoh::ld(&S.klok);
            const unsigned epoch = LdCoh::ld(&S.rng_epoch);
            for (int p = P.p0 + mo * kPBlock + tid; p < P.p0 + P.n; p += Go * kPBlock) noise_one(D, r, p, klok, epoch);
        }
        // pair gather over the replica's particles in cell order, `lanes` lanes
        // each; whole wavefronts run every pass (the groups exchange sums)
        {
            const int lanes = D.pair_lanes, items = P.n * lanes, items_pad = (items + 63) & ~63;
            for (int it = mo * kPBlock + tid; it < items_pad; it += Go * kPBlock)
                pair_gather<R, LdCoh>(D, P.p0 + (it < items ? it : 0) / lanes, it % lanes, lanes, it < items, 31u,
                                      kEvalStep);
        }
    }
    if (!bar_all()) return false;
    // ---- integration: the non-HB members (the HB members wait at the barrier)
    if (!HB) {
        const int n_pad = (P.n + 63) & ~63, mo = m - H, Go = G - H;
        for (int i = mo * kPBlock + tid; i < n_pad; i += Go * kPBlock)
            integrate_one<LdCoh>(D, P.p0 + (i < P.n ? i : 0), i < P.n, (i >> 6) % kMaxSlots);
    }
    if (!bar_all()) return false;
    if (!HB && m == H && wv == 0) commit_replica<LdCoh>(D, r);
    return bar_all();
}

// c_pdev[2 slot]: the state buffers the launch starts from (pos, old, vel,
// f_mb current); c_pdev[2 slot + 1]: the same with the two buffer sets
// exchanged.  Step s uses c_pdev[2 slot + (s & 1)].
template <class R>
__global__ void __launch_bounds__(kPBlock) k_persist(PArgs A) {
    const Dev *Dv = c_pdev + 2 * A.slot;
    extern __shared__ unsigned char lds[];
    __shared__ int sh_ok;
    const int w = blockIdx.x;
    // ---- placement: every workgroup on the XCD of workgroup w % 8
    const int n_rep = Dv[0].n_rep;
    int *gbar = A.bar + 3 * n_rep * kPCtr;
    if (threadIdx.x == 0) {
        const int x = __builtin_amdgcn_s_getreg((20) | (0 << 6) | (15 << 11));   // HW_REG_XCC_ID, bits 0..15
        __hip_atomic_store(A.xcc + w, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (!pbarrier(gbar, gridDim.x, A.status, &sh_ok)) return;
    if (threadIdx.x == 0) {
        const int x = __builtin_amdgcn_s_getreg((20) | (0 << 6) | (15 << 11));
        if (ld_agent_i(A.xcc + (w % kXcds)) != x) atomicOr(A.status, PSTAT_PLACEMENT);
    }
    if (!pbarrier(gbar, 2 * gridDim.x, A.status, &sh_ok)) return;
    if (ld_agent_i(A.status) != 0) return;
    // ---- this workgroup's replica and member index
    const int x = w % kXcds, k = w / kXcds, j = k / A.G, m = k % A.G, r = j * kXcds + x;
    if (j >= A.q || r >= n_rep) return;
    PBars B{A.bar + (3 * r) * kPCtr, A.bar + (3 * r + 1) * kPCtr, A.bar + (3 * r + 2) * kPCtr, 0, 0, 0};
    if (m < A.H) {
        for (int s = 0; s < A.n_steps; ++s)
            if (!persist_step<R, true>(Dv[s & 1], A, r, m, lds, B, &sh_ok)) return;
    } else {
        for (int s = 0; s < A.n_steps; ++s)
            if (!persist_step<R, false>(Dv[s & 1], A, r, m, lds, B, &sh_ok)) return;
    }
}

}  // namespace tis
