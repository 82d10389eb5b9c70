// Kernels of one Langevin step (fast.cpp:6380-6386) for a batch of replicas.
//
//   k_hbenergy    gate, energy and six bead forces of every live HB candidate
//                 (hydrogen_bonds_interactions fast.cpp:2145-2237)
//   k_hbres       valence resolution, grid (hb_groups, replicas): components of the
//                 over-subscribed sites on wavefronts, forces of the live candidates
//                 added and those switched off taken back, energy sums; the
//                 reference-stream noise table (:2239-2350, full_forces draws :2598-2604)
//   k_multibody   stacks, bonds and angles in slices of each replica's beads, forces
//                 summed in LDS (stacking_interactions :1911-2141, polymer_constraints
//                 :2354-2520)
//   k_noise       the step's Philox random forces (full_forces :2590-2625)
//   k_force       pair gather, SR every step and LR when due (SR_/LR_interactions_lists
//                 :1783-1907); evaluations also use it for stacks and backbone terms
//   k_integrate   leap-frog with speed cap, recentring, ion wrap, displacement
//                 reduction (move_rigid_units :2660-2732), then each replica's step
//                 close: rebuild decision (check_shifts :1741-1776)
//   k_bin_count,  cell sort of a rebuild in two parts (charged, uncharged):
//   k_bin_scan,   counts, cell starts, cell-ordered copy (separate_in_cells
//   k_bin_place   :1461-1542)
//   k_list        full SR and LR neighbour lists from the cells within each
//                 particle's list radius (populate_lists :1546-1700)
//   k_hbstatus,   live HB candidates and the per-site tables of the resolution
//   k_hbsites     (populate_lists :1702-1736)
#pragma once
#include "device_math.h"
#include "dual.h"
#include "engine_types.h"

namespace tis {

// ------------------------------------------------------------ block helpers
TIS_DEV int lane_id() { return threadIdx.x & 63; }
TIS_DEV int readlane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
TIS_DEV unsigned long long readlane_u64(unsigned long long v, int l) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), l);
    return ((unsigned long long)hi << 32) | lo;
}
TIS_DEV double readlane_d(double v, int l) { return __longlong_as_double((long long)readlane_u64((unsigned long long)__double_as_longlong(v), l)); }
TIS_DEV double shfl_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __shfl((int)(b & 0xffffffffLL), l, 64), hi = __shfl((int)(b >> 32), l, 64);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

TIS_DEV int wave_id() { return threadIdx.x >> 6; }

// Inclusive prefix sum over the 64 lanes of a wavefront: Hillis-Steele within
// rows of 16 (row_shr 1, 2, 4, 8), then row 0's and rows 0-1's totals carried
// into the rows above them (row_bcast 15 into rows 1 and 3, row_bcast 31 into
// rows 2 and 3).  Lanes with no source read the old operand, 0.
TIS_DEV int wave_incl_scan(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);
    return x;
}


// Ordered exclusive prefix of a per-thread count across the block; *total gets
// the block sum.  scratch must hold blockDim/64 ints.  All threads call.
TIS_DEV int block_excl_sum(int v, int *scratch, int *total) {
    const int lane = lane_id(), w = wave_id(), nw = (blockDim.x + 63) >> 6;
    const int x = wave_incl_scan(v);
    if (lane == 63) scratch[w] = x;
    __syncthreads();
    // every thread adds the wavefront totals itself (independent broadcast reads)
    int before = 0, all = 0;
    for (int i = 0; i < nw; ++i) {
        const int t = scratch[i];
        before += i < w ? t : 0;
        all += t;
    }
    *total = all;
    __syncthreads();   // the caller may reuse scratch at once
    return before + x - v;
}

// Philox key of replica r: (seed low word mixed with the replica, seed high word
// plus the replica's epoch).  The counter carries the step, so draws never
// repeat within an epoch.
TIS_DEV uint32_t philox_key0(const Dev &D, int r) { return (uint32_t)D.seed ^ (0x9E3779B9u * (uint32_t)(r + 1)); }
TIS_DEV uint32_t philox_key1(const Dev &D, const RepState &S) { return (uint32_t)(D.seed >> 32) + S.rng_epoch; }

TIS_DEV double4 d4add(double4 a, double4 b) { return make_double4(a.x + b.x, a.y + b.y, a.z + b.z, 0.0); }

// ------------------------------------------------------------ load policies
// Loads of the arrays a step writes (positions, forces, lists, HB tables, the
// replica state).  LdPlain: an ordinary load; the step's separate kernels only
// read what earlier kernels wrote.  LdCoh: the persistent step (k_persist),
// whose workgroups hand positions, forces and tables to each other between
// the phases of one launch.  A device-scope relaxed atomic load (sc1) misses
// the CU's own L1, which may still hold a line from before the other
// workgroup's store, and reads the L2 that all workgroups of a replica share
// (they run on one XCD, and every store has reached that L2 before the
// barrier that separates the phases).  Wide types load as 8-byte words.
struct LdPlain {
    template <class T>
    TIS_DEV static T ld(const T *p) { return *p; }
};
struct LdCoh {
    template <class T>
    TIS_DEV static T ld(const T *p) {
        return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    TIS_DEV static double4 ld(const double4 *p) {
        const double *q = (const double *)p;
        return make_double4(ld(q), ld(q + 1), ld(q + 2), ld(q + 3));
    }
    TIS_DEV static double2 ld(const double2 *p) {
        const double *q = (const double *)p;
        return make_double2(ld(q), ld(q + 1));
    }
    TIS_DEV static int2 ld(const int2 *p) {
        const long long v = ld((const long long *)p);
        return make_int2((int)(unsigned)v, (int)(v >> 32));
    }
    TIS_DEV static int4 ld(const int4 *p) {
        const long long a = ld((const long long *)p), b = ld((const long long *)p + 1);
        return make_int4((int)(unsigned)a, (int)(a >> 32), (int)(unsigned)b, (int)(b >> 32));
    }
};

// The .w word of a cell-ordered position holds the particle index (low 32 bits)
// and its type word (high 32 bits), so one 32 B load gives a neighbour's
// position, identity and type.
TIS_DEV double cp_pack(int p, int ptype) { return __hiloint2double(ptype, p); }
TIS_DEV int cp_index(double w) { return __double2loint(w); }
TIS_DEV int cp_type(double w) { return __double2hiint(w); }

template <class R>
TIS_DEV void atomic_add_force(double4 *f, int i, V3<R> v) {
    unsafeAtomicAdd(&f[i].x, (double)v.x);
    unsafeAtomicAdd(&f[i].y, (double)v.y);
    unsafeAtomicAdd(&f[i].z, (double)v.z);
}
// Tangent part of a force with derivatives (the backward pass accumulates d F).
TIS_DEV void atomic_add_force(double4 *f, int i, V3<Dual> v) {
    unsafeAtomicAdd(&f[i].x, v.x.d);
    unsafeAtomicAdd(&f[i].y, v.y.d);
    unsafeAtomicAdd(&f[i].z, v.z.d);
}
// Forces of one replica's beads summed in LDS (xyz of bead i - p0 at 3 (i - p0)).
struct LdsForce {
    double *l;
    int p0;
};
template <class R>
TIS_DEV void atomic_add_force(LdsForce f, int i, V3<R> v) {
    double *a = f.l + 3 * (i - f.p0);
    atomicAdd(a, (double)v.x);
    atomicAdd(a + 1, (double)v.y);
    atomicAdd(a + 2, (double)v.z);
}

// Position of particle i as the force routines of precision R see it: the FP64
// position (load policy L), with its tangent from D.jv for derivatives.
template <class R, class L = LdPlain>
struct PtLoad {
    TIS_DEV static double4 at(const Dev &D, int i) { return L::ld(D.pos + i); }
};
template <class L>
struct PtLoad<Dual, L> {
    TIS_DEV static DualPt at(const Dev &D, int i) { return {D.pos[i], D.jv[i]}; }
};

// Block sum, valid in thread 0; scratch holds one double per wavefront.
TIS_DEV double block_sum_double(double v, double *scratch) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    int w = wave_id(), nw = (blockDim.x + 63) >> 6;
    if (lane_id() == 0) scratch[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < nw; ++i) s += scratch[i];
    __syncthreads();
    return s;  // valid in thread 0
}

// ------------------------------------------------------------ pair terms
// WCA force factor (LJ_FORCE, fast.cpp:801-821) times the pair vector.
template <class R>
TIS_DEV R wca_factor(double r2, int ab, const RepParams &P) {
    if (r2 > P.d2lj[ab]) return (R)0;
    R r = sqrt((R)r2);
    R rho = r + (R)kSigLJ - (R)P.dlj[ab];
    R u = (R)kSigLJ / rho;
    R u2 = u * u, u4 = u2 * u2, u8 = u4 * u4, u14 = u8 * u4 * u2;
    return (R)P.flj[ab] * (u14 - u8) * rho / r;
}
// WCA_ENERGY (fast.cpp:749-771).
TIS_DEV double wca_energy(double r2, int ab, const RepParams &P) {
    if (r2 > P.d2lj[ab]) return 0.0;
    if (r2 < P.d2min[ab]) return P.elj[ab] * 1001.0;
    double rho = sqrt(r2) + kSigLJ - P.dlj[ab];
    double u = kSigLJ / rho;
    double u2 = u * u, u4 = u2 * u2, u6 = u4 * u2, u12 = u6 * u6;
    return P.elj[ab] * (u12 - 2.0 * u6 + 1.0);
}

// ------------------------------------------------------------ cell sort
// The cell of position x (separate_in_cells, fast.cpp:1510-1530): each
// coordinate wrapped once into the box, cell ceil(t / lc) - 1 clamped to the
// grid, the quotient from the reciprocal, and from the division itself when it
// lies within 1e-9 of an integer (where the two may round apart).  *outside is
// set when a wrapped coordinate still lies outside [0, L] (its particle is then
// not inside its cell, which the cropped list search relies on).
TIS_DEV int cell_index(const double x[3], const double bL[3], const double blc[3], const double binv[3],
                       const int bM[3], bool *outside) {
    int ci[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double t = x[a];
        if (t < 0) t += bL[a];
        else if (t > bL[a]) t -= bL[a];
        *outside |= !(t >= 0.0 && t <= bL[a]);
        double qt = t * binv[a];
        if (fabs(qt - rint(qt)) < 1e-9 * (1.0 + fabs(qt))) qt = t / blc[a];
        const int c = (int)ceil(qt) - 1;
        ci[a] = c < 0 ? 0 : (c >= bM[a] ? bM[a] - 1 : c);
    }
    return (ci[0] * bM[1] + ci[1]) * bM[2] + ci[2];
}
// The cell order holds each replica's charged particles first, then its
// uncharged ones, each part sorted by cell: the sort runs over 2 nc "virtual"
// cells, cell c of the charged part being c and of the uncharged part nc + c,
// so the long-range search of a charged particle (charged partners only) walks
// the charged part alone.
TIS_DEV int virtual_cell(int cell, int nc, int ptype) { return ((ptype >> 8) & 1) ? cell : nc + cell; }

// Counting sort of replica r's particles into virtual cells by one workgroup
// (any multiple of 64 threads up to 1024); cnt holds 2 kMaxCells ints, scratch
// 20, ord ord_cap 16-bit entries (the cell order, when the replica has at most
// ord_cap particles).
constexpr int kBinBatch = 8;
template <class L>
TIS_DEV void bin_replica(const Dev &D, int r, int *cnt, int *scratch, unsigned short *ord, int ord_cap) {
    RepState &S = D.rs[r];
    const RepParams &P = D.rp[r];
    const int M0 = P.M[0], M1 = P.M[1], M2 = P.M[2], nc = M0 * M1 * M2;
    // the replica's box and cell sizes in registers: the stores below would
    // otherwise make the compiler load them again for every particle
    const double bL[3] = {P.L[0], P.L[1], P.L[2]}, blc[3] = {P.lc[0], P.lc[1], P.lc[2]};
    const double binv[3] = {1.0 / blc[0], 1.0 / blc[1], 1.0 / blc[2]};
    const int bM[3] = {M0, M1, M2};
    const int p0 = P.p0, n = P.n, cell0 = P.cell0, nv = 2 * nc;
    for (int c = threadIdx.x; c < nv; c += blockDim.x) cnt[c] = 0;
    __syncthreads();
    // kBinBatch particles per thread at a time, their loads issued together
    for (int i0 = threadIdx.x; i0 < n; i0 += kBinBatch * blockDim.x) {
        double4 xb[kBinBatch];
        int tb[kBinBatch];
#pragma unroll
        for (int k = 0; k < kBinBatch; ++k) {
            const int i = i0 + k * blockDim.x;
            xb[k] = i < n ? L::ld(D.pos + p0 + i) : make_double4(0, 0, 0, 0);
            tb[k] = i < n ? D.ptype[p0 + i] : 0;
        }
#pragma unroll
        for (int k = 0; k < kBinBatch; ++k) {
            const int i = i0 + k * blockDim.x;
            if (i >= n) break;
            const int p = p0 + i;
            const double xs[3] = {xb[k].x, xb[k].y, xb[k].z};
            bool outside = false;
            const int cell = virtual_cell(cell_index(xs, bL, blc, binv, bM, &outside), nc, tb[k]);
            if (outside) atomicOr(&S.crop_off, 1);
            D.cell_of[p] = cell;
            D.cell_rank[p] = atomicAdd(&cnt[cell], 1);
        }
    }
    __syncthreads();
    // exclusive scan of the cell counts: contiguous chunk per thread, then block scan
    const int per = (nv + blockDim.x - 1) / blockDim.x;
    const int c0 = threadIdx.x * per;
    int local = 0;
    for (int c = c0; c < c0 + per && c < nv; ++c) local += cnt[c];
    int total;
    int run = block_excl_sum(local, scratch, &total);
    for (int c = c0; c < c0 + per && c < nv; ++c) {
        int v = cnt[c];
        cnt[c] = run;
        D.cell_start[cell0 + c] = run;
        run += v;
    }
    if (threadIdx.x == 0) D.cell_start[cell0 + nv] = total;
    __syncthreads();
    if (n <= ord_cap) {
        // the cell order in LDS first (ord[q - p0] = p - p0), then the cell-ordered
        // arrays written in cell order: every global store of the sort lands on
        // consecutive addresses (scattered partial-line stores leave the kernel's
        // end-of-kernel cache write-back far slower)
        for (int i0 = threadIdx.x; i0 < n; i0 += kBinBatch * blockDim.x) {
            int cb[kBinBatch], rb[kBinBatch];
#pragma unroll
            for (int k = 0; k < kBinBatch; ++k) {
                const int i = i0 + k * blockDim.x, p = p0 + (i < n ? i : 0);
                cb[k] = L::ld(D.cell_of + p);
                rb[k] = L::ld(D.cell_rank + p);
            }
#pragma unroll
            for (int k = 0; k < kBinBatch; ++k) {
                const int i = i0 + k * blockDim.x;
                if (i >= n) break;
                const int ql = cnt[cb[k]] + rb[k];
                ord[ql] = (unsigned short)i;
                D.crank[p0 + i] = p0 + ql;
            }
        }
        __syncthreads();
        for (int q0 = threadIdx.x; q0 < n; q0 += kBinBatch * blockDim.x) {
            int pb[kBinBatch], tb[kBinBatch];
            double4 xb[kBinBatch];
#pragma unroll
            for (int k = 0; k < kBinBatch; ++k) {
                const int ql = q0 + k * blockDim.x;
                pb[k] = p0 + (ql < n ? (int)ord[ql] : 0);
                tb[k] = D.ptype[pb[k]];
                xb[k] = L::ld(D.pos + pb[k]);
            }
#pragma unroll
            for (int k = 0; k < kBinBatch; ++k) {
                const int ql = q0 + k * blockDim.x;
                if (ql >= n) break;
                D.sorted[p0 + ql] = pb[k];
                D.cpos[p0 + ql] = make_double4(xb[k].x, xb[k].y, xb[k].z, cp_pack(pb[k], tb[k]));
            }
        }
        if (threadIdx.x == 0) {
            S.max_sr = 0;
            S.max_lr = 0;
        }
        return;
    }
    for (int i0 = threadIdx.x; i0 < n; i0 += kBinBatch * blockDim.x) {
        int cb[kBinBatch], rb[kBinBatch], tb[kBinBatch];
        double4 xb[kBinBatch];
#pragma unroll
        for (int k = 0; k < kBinBatch; ++k) {
            const int i = i0 + k * blockDim.x, p = p0 + (i < n ? i : 0);
            cb[k] = L::ld(D.cell_of + p);
            rb[k] = L::ld(D.cell_rank + p);
            tb[k] = D.ptype[p];
            xb[k] = L::ld(D.pos + p);
        }
#pragma unroll
        for (int k = 0; k < kBinBatch; ++k) {
            const int i = i0 + k * blockDim.x;
            if (i >= n) break;
            const int p = p0 + i;
            const int q = p0 + cnt[cb[k]] + rb[k];
            D.sorted[q] = p;
            D.crank[p] = q;
            D.cpos[q] = make_double4(xb[k].x, xb[k].y, xb[k].z, cp_pack(p, tb[k]));
        }
    }
    if (threadIdx.x == 0) {
        S.max_sr = 0;
        S.max_lr = 0;
    }
}
// The same sort in three launches over many workgroups (a replica's particles
// need not wait on one CU).  k_bin_count, grid (ceil(n_max / 1024), replicas):
// each workgroup ranks its 1024 particles within their cells in LDS, then
// reserves each touched cell's share of the replica's count with one global
// atomic per cell (returning the share's base): rank = base + rank in the share.
// k_bin_scan, one workgroup per replica: the cell starts (an exclusive scan of
// the counts, which it clears for the next rebuild).  k_bin_place, grid as
// k_bin_count: cell-order index = cell start + rank, and the cell-ordered
// arrays.  Only replicas whose trigger fired do anything.
__global__ void __launch_bounds__(1024) k_bin_count(Dev D) {
    const int r = blockIdx.y;
    if (!D.rs[r].need_rebuild) return;
    const RepParams &P = D.rp[r];
    const int n = P.n, p0 = P.p0, cell0 = P.cell0;
    if ((int)(blockIdx.x * 1024) >= n) return;
    const int bM[3] = {P.M[0], P.M[1], P.M[2]}, nc = bM[0] * bM[1] * bM[2], nv = 2 * nc;
    const double bL[3] = {P.L[0], P.L[1], P.L[2]}, blc[3] = {P.lc[0], P.lc[1], P.lc[2]};
    const double binv[3] = {1.0 / blc[0], 1.0 / blc[1], 1.0 / blc[2]};
    __shared__ int lcnt[2 * kMaxCells];
    const int i = blockIdx.x * 1024 + threadIdx.x;
    double4 x = make_double4(0, 0, 0, 0);
    int pt = 0;
    if (i < n) {
        x = D.pos[p0 + i];
        pt = D.ptype[p0 + i];
    }
    for (int c = threadIdx.x; c < nv; c += 1024) lcnt[c] = 0;
    __syncthreads();
    int cell = 0, lrank = 0;
    if (i < n) {
        const double xs[3] = {x.x, x.y, x.z};
        bool outside = false;
        cell = virtual_cell(cell_index(xs, bL, blc, binv, bM, &outside), nc, pt);
        lrank = atomicAdd(&lcnt[cell], 1);
        if (outside) atomicOr(&D.rs[r].crop_off, 1);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < nv; c += 1024) {
        const int k = lcnt[c];
        if (k > 0) lcnt[c] = atomicAdd(&D.cell_cnt[cell0 + c], k);
    }
    __syncthreads();
    if (i < n) {
        D.cell_of[p0 + i] = cell;
        D.cell_rank[p0 + i] = lcnt[cell] + lrank;
    }
}
__global__ void __launch_bounds__(1024) k_bin_scan(Dev D) {
    const int r = blockIdx.x;
    RepState &S = D.rs[r];
    if (!S.need_rebuild) return;
    const RepParams &P = D.rp[r];
    const int nc = 2 * P.M[0] * P.M[1] * P.M[2], cell0 = P.cell0;   // virtual cells
    __shared__ int scratch[20];
    __shared__ int sc[2 * kMaxCells];
    // the counts staged through LDS so that global loads and stores are
    // consecutive across lanes (a thread's contiguous run of cells would
    // scatter them)
    constexpr int kPer = (2 * kMaxCells + 1023) / 1024;
    int v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int c = threadIdx.x + 1024 * k;
        v[k] = c < nc ? D.cell_cnt[cell0 + c] : 0;
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int c = threadIdx.x + 1024 * k;
        if (c < nc) sc[c] = v[k];
    }
    __syncthreads();
    const int per = (nc + 1023) / 1024, c0 = threadIdx.x * per;
    int local = 0;
    for (int c = c0; c < c0 + per && c < nc; ++c) local += sc[c];
    int total;
    int run = block_excl_sum(local, scratch, &total);
    for (int c = c0; c < c0 + per && c < nc; ++c) {
        const int x = sc[c];
        sc[c] = run;
        run += x;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int c = threadIdx.x + 1024 * k;
        if (c < nc) {
            D.cell_start[cell0 + c] = sc[c];
            D.cell_cnt[cell0 + c] = 0;   // cleared for the next rebuild
        }
    }
    if (threadIdx.x == 0) {
        D.cell_start[cell0 + nc] = total;
        S.max_sr = 0;
        S.max_lr = 0;
    }
}
__global__ void __launch_bounds__(1024) k_bin_place(Dev D) {
    const int r = blockIdx.y;
    if (!D.rs[r].need_rebuild) return;
    const RepParams &P = D.rp[r];
    const int n = P.n, p0 = P.p0, cell0 = P.cell0;
    const int i = blockIdx.x * 1024 + threadIdx.x;
    if (i >= n) return;
    const int p = p0 + i;
    const int cell = D.cell_of[p], rank = D.cell_rank[p], pt = D.ptype[p];
    const double4 x = D.pos[p];
    const int q = p0 + D.cell_start[cell0 + cell] + rank;
    D.sorted[q] = p;
    D.crank[p] = q;
    D.cpos[q] = make_double4(x.x, x.y, x.z, cp_pack(p, pt));
}

// ------------------------------------------------------------ neighbour cells
// Cells are at least half the long-range list radius wide, so the particles
// within r_lr of a particle in cell (c0, c1, c2) lie in cells c +- 2 along each
// axis (h = 2), and, when every cell is at least r_sr wide, those within r_sr in
// cells c +- 1 (h = 1: particles that get no long-range list).  Cells are
// numbered with z fastest, so the 2h + 1 cells c2 - h .. c2 + h of one (x, y)
// column are contiguous in cell order (two pieces where the range wraps the box):
// the neighbourhood is at most 5 x 5 columns x 2 pieces = 50 contiguous ranges of
// particles, one per lane: lane l < *nseg gets its range's first cell-order index
// and length (0 for an empty piece).  An axis with fewer than 2h + 1 cells
// contributes each of its cells once.  With rc2 >= 0 the ranges are cut to the
// cells that can hold a point within sqrt(rc2) of xi, the particle's position
// (every particle lies in the cell the sort gave it unless the replica's
// crop_off is set; rc2 < 0 takes the whole neighbourhood).  seg: the charged
// (0) or uncharged (1) part of the cell order.  The ranges go to lanes lane0,
// lane0 + 1, ... (the other lanes keep theirs, so two calls can fill one
// wavefront); returns the lane after the last.
template <class L>
TIS_DEV int neighbour_cells(const Dev &D, const RepParams &P, int cell, int h, int seg, const double4 &xi,
                            double rc2, int lane0, int *sstart, int *slen) {
    const int M0 = P.M[0], M1 = P.M[1], M2 = P.M[2];
    const int c2 = cell % M2, c1 = (cell / M2) % M1, c0 = cell / (M1 * M2);
    const int w = 2 * h + 1;
    const int wx = M0 < w ? M0 : w, wy = M1 < w ? M1 : w;
    const int lane = threadIdx.x & 63;
    const int col = (lane - lane0) >> 1, piece = (lane - lane0) & 1;
    if (lane >= lane0 && col < wx * wy) {
        *sstart = 0;
        *slen = 0;
        const int ix = col / wy, iy = col % wy;
        // cropping (rc2 >= 0): a neighbour cell u along an axis of at least 2h + 1
        // cells lies at least gap(u) = max(0, u lc - t, t - (u + 1) lc) from the
        // particle (t: its coordinate wrapped once, as the cell sort wraps it; u
        // unwrapped, so this image is the nearest); columns whose x and y gaps
        // exceed the radius are skipped and the z range is cut to the cells
        // within the radius left
        const bool crop = rc2 >= 0.0;
        double rem = rc2;
        auto gap = [&](int a, double x, int u) {
            double t = x;
            if (t < 0) t += P.L[a];
            else if (t > P.L[a]) t -= P.L[a];
            const double lc = P.lc[a];
            return fmax(0.0, fmax(u * lc - t, t - (u + 1) * lc));
        };
        int n0 = ix, n1 = iy;
        if (M0 >= w) {
            n0 = c0 - h + ix;
            if (crop) { const double g = gap(0, xi.x, n0); rem -= g * g; }
            n0 += n0 < 0 ? M0 : 0; n0 -= n0 >= M0 ? M0 : 0;
        }
        if (M1 >= w) {
            n1 = c1 - h + iy;
            if (crop) { const double g = gap(1, xi.y, n1); rem -= g * g; }
            n1 += n1 < 0 ? M1 : 0; n1 -= n1 >= M1 ? M1 : 0;
        }
        int zlo = 0, zhi = M2 - 1;           // this piece's z range (zhi < zlo: empty)
        if (crop && rem < 0.0) {
            zhi = -1;
        } else if (M2 >= w) {
            int lo = c2 - h, hi = c2 + h;
            if (crop) {
                double t = xi.z;
                if (t < 0) t += P.L[2];
                else if (t > P.L[2]) t -= P.L[2];
                const double sz = sqrt(rem), lc = P.lc[2];
                lo = max(lo, min(c2, (int)ceil((t - sz) / lc) - 1));
                hi = min(hi, max(c2, (int)floor((t + sz) / lc)));
            }
            if (lo < 0) {
                zlo = piece ? lo + M2 : 0;
                zhi = piece ? M2 - 1 : hi;
            } else if (hi >= M2) {
                zlo = piece ? 0 : lo;
                zhi = piece ? hi - M2 : M2 - 1;
            } else {
                zlo = lo;
                zhi = piece ? lo - 1 : hi;
            }
        } else if (piece) {
            zhi = -1;
        }
        if (zhi >= zlo) {
            const int base = P.cell0 + seg * M0 * M1 * M2 + (n0 * M1 + n1) * M2;
            *sstart = L::ld(D.cell_start + base + zlo);
            *slen = L::ld(D.cell_start + base + zhi + 1) - *sstart;
        }
    }
    return lane0 + 2 * wx * wy;
}

// Visits every particle of the ranges that lanes 0 .. ncell - 1 hold (from
// neighbour_cells), 64 per call of visit(j, x, q) (particle j, its cell-ordered
// position x at global cell-order index q), lanes past the last candidate
// getting j = -1.  The cells' particles are taken
// as one sequence (cell c covers positions [excl_c, incl_c) of it, prefix sums
// over the lanes that hold the cells), so every chunk of 64 lanes is full
// whatever the cell occupancies; a lane finds its cell by a scalar walk over the
// few cells that overlap its chunk.  Four chunks are in flight: the load of a
// chunk is issued three chunks before it is tested (the loop is unrolled four
// times so that each chunk keeps its registers until it is tested).  Every
// lane loads in every chunk (a lane without a candidate reads the replica's
// first entry and is marked by q = -1), so that the loads form a fixed sequence
// the compiler can count in flight instead of waiting for all of them.
template <class L, class F>
TIS_DEV void for_cell_ranges_wave(const Dev &D, const RepParams &P, int ncell, int cstart, int clen, F &&visit) {
    const int lane = threadIdx.x & 63;
    const int incl = wave_incl_scan(clen);
    const int total = readlane_i(incl, 63);
    if (total == 0) return;
    const int p0 = P.p0;
    const double4 *cp = D.cpos + p0;
    int c = 0;   // first cell whose range ends after the chunk start (uniform)
    auto load = [&](int t0, double4 *x, int *q) {
        int addr = -1;
        if (t0 < total) {
            while (readlane_i(incl, c) <= t0) ++c;
            const int t = t0 + lane;
            for (int cc = c; cc < ncell; ++cc) {
                const int in = readlane_i(incl, cc), ex = in - readlane_i(clen, cc);
                if (ex >= t0 + 64) break;
                if (t >= ex && t < in) addr = readlane_i(cstart, cc) + (t - ex);
            }
        }
        *x = L::ld(cp + (addr >= 0 ? addr : 0));
        *q = addr >= 0 ? p0 + addr : -1;
    };
    auto run = [&](const double4 &x, int q) { visit(q >= 0 ? cp_index(x.w) : -1, x, q); };
    double4 x0, x1, x2, x3;
    int q0, q1, q2, q3;
    load(0, &x0, &q0);
    load(64, &x1, &q1);
    load(128, &x2, &q2);
    load(192, &x3, &q3);
    for (int t0 = 0;; t0 += 256) {
        run(x0, q0);
        if (t0 + 64 >= total) break;
        load(t0 + 256, &x0, &q0);
        run(x1, q1);
        if (t0 + 128 >= total) break;
        load(t0 + 320, &x1, &q1);
        run(x2, q2);
        if (t0 + 192 >= total) break;
        load(t0 + 384, &x2, &q2);
        run(x3, q3);
        if (t0 + 256 >= total) break;
        load(t0 + 448, &x3, &q3);
    }
}

// ------------------------------------------------------------ k_list
// One wavefront per particle: the lanes test 64 candidates at a time, four
// chunks in flight.  Every particle owns a fixed slot of the pool (cap_sr SR
// then cap_lr LR entries), so there is no counting pass and no shared cursor.
// The grid is fixed (kListWaves wavefronts) and strides over the particles of
// the replicas whose displacement trigger fired; the others cost one flag read.
// MODE 0: record the largest per-particle counts only (sizes the slots), MODE 1:
// build.  A chunk's stores sit behind a wave-uniform test of its ballot (most
// chunks hold no short-range partner, and uncharged particles no long-range one).
constexpr int kListWaves = 4096;
// Lists of replica r's particles, wavefront w0 of nw taking particles w0, w0 + nw, ...
template <class L, int MODE>
TIS_DEV void list_replica(const Dev &D, int r, int w0, int nw) {
    const int lane = threadIdx.x & 63;
    {
        RepState &S = D.rs[r];
        const RepParams &P = D.rp[r];
        // the replica's box, radii and slot sizes in registers: the list stores
        // below would otherwise make the compiler load them again for every chunk
        const double bL[3] = {P.L[0], P.L[1], P.L[2]}, bhL[3] = {P.hL[0], P.hL[1], P.hL[2]};
        const double rsr2 = P.rsr2, rlr2 = P.rlr2;
        const int cap_sr = P.cap_sr, cap_lr = P.cap_lr, p0 = P.p0, n = P.n, sr_h = P.sr_h;
        const int nc = P.M[0] * P.M[1] * P.M[2];
        const long long pool0 = P.pool0;
        // search radii of the cell cropping, 1e-3 A beyond the list radii (each
        // particle lies within 1e-12 A of its cell); -1: no cropping
        const bool crop = !L::ld(&S.crop_off);
        const double rc_sr = sqrt(rsr2) + 1e-3, rc_lr = sqrt(rlr2) + 1e-3;
        const double rc2_sr = crop ? rc_sr * rc_sr : -1.0, rc2_lr = crop ? rc_lr * rc_lr : -1.0;
        const long long stride = (long long)cap_sr + cap_lr;
        for (int t = w0; t < n; t += nw) {
            const double4 xi = L::ld(D.cpos + p0 + t);
            const int p = cp_index(xi.w);
            const bool chg_i = (cp_type(xi.w) >> 8) & 1;
            const int vcell = L::ld(D.cell_of + p), cell = vcell >= nc ? vcell - nc : vcell;
            // entries are cell-order indices q of the partners; bonded / 1-3
            // partners of p are stored as ~q, so the step path leaves out their WCA term
            int ex[kExcl];
#pragma unroll
            for (int e = 0; e < kExcl; ++e) ex[e] = D.excl[kExcl * p + e];
            int *out_sr = D.pool + pool0 + (long long)(p - p0) * stride;
            int *out_lr = out_sr + cap_sr;
            int a = 0, b = 0;
            auto visit = [&](int j, const double4 &xj, int qj) {
                const double r2 = min_image_r2(xi, xj, bL, bhL);
                const bool cand = j >= 0 && j != p;
                const bool isr = cand && r2 <= rsr2;
                const bool ilr = cand && !isr && chg_i && r2 <= rlr2 && ((cp_type(xj.w) >> 8) & 1);
                const unsigned long long ms = __ballot(isr), ml = __ballot(ilr);
                if (MODE == 1) {
                    const unsigned long long below = (1ull << lane) - 1ull;
                    if (ms) {   // most chunks hold no short-range partner
                        bool is_ex = false;
#pragma unroll
                        for (int e = 0; e < kExcl; ++e) is_ex |= ex[e] == j;
                        const int ka = a + __popcll(ms & below);
                        if (isr && ka < cap_sr) out_sr[ka] = is_ex ? ~qj : qj;
                    }
                    if (ml) {
                        const int kb = b + __popcll(ml & below);
                        if (ilr && kb < cap_lr) out_lr[kb] = qj;
                    }
                }
                a += __popcll(ms);
                b += __popcll(ml);
            };
            // a charged particle: the charged part within r_lr (its long-range
            // partners and its charged short-range ones), then the uncharged part
            // within r_sr; an uncharged particle: both parts within r_sr (one
            // walk when both fit a wavefront's lanes)
            int cs = 0, cl = 0;
            if (chg_i) {
                int nr = neighbour_cells<L>(D, P, cell, 2, 0, xi, rc2_lr, 0, &cs, &cl);
                for_cell_ranges_wave<L>(D, P, nr, cs, cl, visit);
                cs = 0;
                cl = 0;
                nr = neighbour_cells<L>(D, P, cell, sr_h, 1, xi, rc2_sr, 0, &cs, &cl);
                for_cell_ranges_wave<L>(D, P, nr, cs, cl, visit);
            } else if (sr_h == 1) {
                int nr = neighbour_cells<L>(D, P, cell, 1, 0, xi, rc2_sr, 0, &cs, &cl);
                nr = neighbour_cells<L>(D, P, cell, 1, 1, xi, rc2_sr, nr, &cs, &cl);
                for_cell_ranges_wave<L>(D, P, nr, cs, cl, visit);
            } else {
                for (int seg = 0; seg < 2; ++seg) {
                    cs = 0;
                    cl = 0;
                    const int nr = neighbour_cells<L>(D, P, cell, sr_h, seg, xi, rc2_sr, 0, &cs, &cl);
                    for_cell_ranges_wave<L>(D, P, nr, cs, cl, visit);
                }
            }
            if (lane == 0) {
                if (MODE == 0) {
                    atomicMax(&S.max_sr, a);
                    atomicMax(&S.max_lr, b);
                } else {
                    if (a > cap_sr || b > cap_lr) atomicOr(&S.overflow, 1);
                    D.nb_start[p] = (int)((p - p0) * stride);
                    D.nb_nsr[p] = min(a, cap_sr);
                    D.nb_nlr[p] = min(b, cap_lr);
                    D.nb_meta[p0 + t] = make_int4((int)((pool0 + (p - p0) * stride) >> 4), min(a, cap_sr),
                                                  min(b, cap_lr), r);
                }
            }
        }
    }
}
__global__ void __launch_bounds__(kBlock) k_list(Dev D, int mode) {
    const int w0 = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    // the rebuild flags of 64 replicas at a time, one per lane, loaded at once
    const int lane = threadIdx.x & 63;
    for (int r0 = 0; r0 < D.n_rep; r0 += 64) {
        unsigned long long due = __ballot(r0 + lane < D.n_rep && D.rs[r0 + lane].need_rebuild);
        while (due) {
            const int r = r0 + __builtin_ctzll(due);
            due &= due - 1ull;
            if (mode == 1) list_replica<LdPlain, 1>(D, r, w0, nw);
            else list_replica<LdPlain, 0>(D, r, w0, nw);
        }
    }
}

// ------------------------------------------------------------ HB rebuild
// One buffer of the HB rebuild tables (the arrays hold two, tab_* elements apart).
struct HbTab {
    int *act_hb, *asite_val, *asite_off, *asite_list, *ahb_inc_off, *ahb_inc;
    int4 *act_rec;   // per listed candidate: (id, beads 0-2), (beads 3-5, 0)
};
TIS_DEV HbTab hb_tab(const Dev &D, int b) {
    return {D.act_hb + b * D.tab_hb,      D.asite_val + b * D.tab_site,    D.asite_off + b * D.tab_site,
            D.asite_list + b * D.tab_inc, D.ahb_inc_off + b * D.tab_hb, D.ahb_inc + b * D.tab_inc,
            D.act_rec + 2 * b * D.tab_hb};
}

// k_hbstatus: grid-stride over the candidates of the replicas being rebuilt.
// A candidate is live iff its (k11, k21) bead pair is an SR pair, decided with
// the list's own FP64 distance test (populate_lists, fast.cpp:1702-1727).  Live
// candidates are counted at each of their sites; k_hbsites clears the counters
// after reading them.
constexpr int kStatusBlock = 1024;
constexpr int kMaxHbChunks = 64;   // candidates per replica <= kMaxHbChunks * kStatusBlock
// One chunk of kStatusBlock candidates of replica r (a workgroup of
// kStatusBlock threads).
template <class L>
TIS_DEV void hbstatus_chunk(const Dev &D, int lag, int r, int chunk, int *scratch) {
    const RepParams &P = D.rp[r];
    const int k = chunk * kStatusBlock + threadIdx.x;
    bool live = false;
    if (k < P.nhb) {
        const int g = P.hb0 + k;
        const int b1 = D.hb_b[6 * g + 1], b4 = D.hb_b[6 * g + 4];
        live = min_image_r2(L::ld(D.pos + b1), L::ld(D.pos + b4), P.L, P.hL) <= P.rsr2;
        D.hb_status[g] = live ? 0 : -1;
        if (!lag) D.hb_out[g] = live ? 0 : -1;   // lagged: the resolution running beside owns hb_out
        if (live)
            for (int q = D.hb_inc_off[g]; q < D.hb_inc_off[g + 1]; ++q) atomicAdd(&D.site_cnt[D.hb_inc[q]], 1);
    }
    // rank among the chunk's live candidates (candidate order); k_hbsites adds the
    // chunk offsets
    int tot;
    const int rank = block_excl_sum(live, scratch, &tot);
    if (live) D.act_tmp[P.hb0 + chunk * kStatusBlock + rank] = P.hb0 + k;   // the chunk's live ids, packed
    if (threadIdx.x == 0) D.chunk_cnt[kMaxHbChunks * r + chunk] = tot;
}
__global__ void __launch_bounds__(kStatusBlock) k_hbstatus(Dev D, int lag) {
    const int r = blockIdx.y;
    if (!D.rs[r].need_rebuild) return;
    if ((int)(blockIdx.x * kStatusBlock) >= D.rp[r].nhb) return;
    __shared__ int scratch[20];
    hbstatus_chunk<LdPlain>(D, lag, r, blockIdx.x, scratch);
}

// k_hbsites: one workgroup per rebuilt replica.  Active sites (with at least one
// live candidate) in site order, with their valences, each with its live candidates in the order of
// the static site list HB_ATOM (restored from the precomputed rank of every
// incidence), and every live candidate's incidences as active-site indices.
// Dynamic LDS: 2 * nsite ints.  lag: write the buffer the resolution is not
// reading (it runs beside this kernel) and leave the switch to the step's close.
// LDS of one workgroup of hbsites_replica: 2 nsite ints (site map, fill
// counts), the chunk offsets, the site segment offsets and the segments.
struct HbSitesLds {
    int *smem, *scratch, *chunk_off, *aoff_l, *seg_l;
};
template <class L>
TIS_DEV void hbsites_replica(const Dev &D, int lag, int r, HbSitesLds W) {
    RepState &S = D.rs[r];
    const RepParams &Pg = D.rp[r];
    // the replica's table offsets in registers: the stores below would otherwise
    // make the compiler load them again in every loop
    struct {
        int hb0, nhb, site0, nsite, inc0;
    } const P = {Pg.hb0, Pg.nhb, Pg.site0, Pg.nsite, Pg.inc0};
    const int tab = L::ld(&S.tab);
    const HbTab T = hb_tab(D, lag ? tab ^ 1 : tab);
    int *site_map = W.smem, *fill = W.smem + P.nsite;
    int *scratch = W.scratch, *chunk_off = W.chunk_off;
    int *aoff_l = W.aoff_l;   // site segment offsets
    int *seg_l = W.seg_l;     // (static rank << 16 | live index) per site segment
    // live candidates in candidate order: chunk offsets, then chunk-local ranks
    const int nchunk = (P.nhb + kStatusBlock - 1) / kStatusBlock;
    if (threadIdx.x < 64) {
        const int c = threadIdx.x < nchunk ? L::ld(D.chunk_cnt + kMaxHbChunks * r + threadIdx.x) : 0;
        const int incl = wave_incl_scan(c);
        chunk_off[threadIdx.x] = incl - c;
        if (threadIdx.x == 63) chunk_off[kMaxHbChunks] = incl;
    }
    __syncthreads();
    const int n_act = chunk_off[kMaxHbChunks];
    if (n_act <= kActMax) {
        // live index a lies in chunk c with chunk_off[c] <= a < chunk_off[c + 1]
        // (binary search; empty chunks share their offset with the next one)
        for (int a = threadIdx.x; a < n_act; a += blockDim.x) {
            int lo = 0, hi = nchunk;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (chunk_off[mid] <= a) lo = mid;
                else hi = mid;
            }
            const int g = L::ld(D.act_tmp + P.hb0 + lo * kStatusBlock + (a - chunk_off[lo]));
            T.act_hb[P.hb0 + a] = g;
            const int *hb = D.hb_b + 6 * g;
            T.act_rec[2 * (P.hb0 + a)] = make_int4(g, hb[0], hb[1], hb[2]);
            T.act_rec[2 * (P.hb0 + a) + 1] = make_int4(hb[3], hb[4], hb[5], 0);
        }
    }
    __syncthreads();
    if (n_act > kActMax) {
        for (int q = threadIdx.x; q < P.nsite; q += blockDim.x) D.site_cnt[P.site0 + q] = 0;
        if (threadIdx.x == 0) atomicOr(&S.overflow, 2);
        return;
    }
    // active sites in site order: each thread a contiguous range of sites, one
    // scan each for the active-site ranks and the incidence offsets
    const int per = (P.nsite + blockDim.x - 1) / blockDim.x;
    const int s_lo = min(P.nsite, (int)threadIdx.x * per), s_hi = min(P.nsite, s_lo + per);
    int my_as = 0, my_inc = 0;
    for (int s = s_lo; s < s_hi; ++s) {
        const int c = L::ld(D.site_cnt + P.site0 + s);
        my_as += c > 0;
        my_inc += c;
    }
    int n_as, n_inc;
    int rank = block_excl_sum(my_as, scratch, &n_as);
    int off = block_excl_sum(my_inc, scratch, &n_inc);
    for (int s = s_lo; s < s_hi; ++s) {
        const int cnt = L::ld(D.site_cnt + P.site0 + s);
        D.site_cnt[P.site0 + s] = 0;   // cleared for the next build
        fill[s] = 0;
        if (cnt > 0) {
            site_map[s] = rank;
            if (rank < kASiteMax) {
                T.asite_val[P.site0 + rank] = D.site_val[P.site0 + s];
                T.asite_off[P.site0 + r + rank] = off;
                aoff_l[rank] = off;
            }
            ++rank;
            off += cnt;
        } else {
            site_map[s] = -1;
        }
    }
    if (threadIdx.x == 0 && n_as <= kASiteMax) {
        T.asite_off[P.site0 + r + n_as] = n_inc;
        aoff_l[n_as] = n_inc;
    }
    __syncthreads();
    if (n_as > kASiteMax || n_inc > kAIncMax) {
        if (threadIdx.x == 0) atomicOr(&S.overflow, 2);
        return;
    }
    // scatter (rank << 16 | live index) into each active site's segment
    int n_ai = 0;
    for (int a0 = 0; a0 < n_act; a0 += blockDim.x) {
        const int a = a0 + threadIdx.x;
        int c = 0, g = 0;
        if (a < n_act) {
            g = L::ld(T.act_hb + P.hb0 + a);
            c = D.hb_inc_off[g + 1] - D.hb_inc_off[g];
        }
        int tot;
        const int off = block_excl_sum(c, scratch, &tot);
        if (a < n_act) {
            T.ahb_inc_off[P.hb0 + r + a] = n_ai + off;
            for (int q = 0; q < c; ++q) {
                const int e = D.hb_inc_off[g] + q;
                const int ai = site_map[D.hb_inc[e] - P.site0];
                T.ahb_inc[P.inc0 + n_ai + off + q] = ai;
                const int slot = atomicAdd(&fill[D.hb_inc[e] - P.site0], 1);
                seg_l[aoff_l[ai] + slot] = (D.inc_rank[e] << 16) | a;
            }
        }
        n_ai += tot;
    }
    if (threadIdx.x == 0) T.ahb_inc_off[P.hb0 + r + n_act] = n_ai;
    __syncthreads();
    // order each site's segment by static rank (in LDS), keep the live index
    for (int ai = threadIdx.x; ai < n_as; ai += blockDim.x) {
        int *seg = seg_l + aoff_l[ai];
        const int len = aoff_l[ai + 1] - aoff_l[ai];
        for (int i = 1; i < len; ++i) {
            const int v = seg[i];
            int j = i - 1;
            while (j >= 0 && seg[j] > v) { seg[j + 1] = seg[j]; --j; }
            seg[j + 1] = v;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n_inc; i += blockDim.x) T.asite_list[P.inc0 + i] = seg_l[i] & 0xffff;
    if (threadIdx.x == 0) {
        if (lag) {
            S.nx_act = n_act;
            S.nx_asite = n_as;
            S.nx_ainc = n_inc;
            S.tab_pending = 1;
        } else {
            S.n_act = n_act;
            S.n_asite = n_as;
            S.n_ainc = n_inc;
        }
    }
}
__global__ void __launch_bounds__(1024) k_hbsites(Dev D, int lag) {
    const int r = blockIdx.x;
    if (!D.rs[r].need_rebuild) return;
    extern __shared__ int smem[];
    __shared__ int scratch[20];
    __shared__ int chunk_off[kMaxHbChunks + 1];
    __shared__ int aoff_l[kASiteMax + 1];
    __shared__ int seg_l[kAIncMax];
    hbsites_replica<LdPlain>(D, lag, r, HbSitesLds{smem, scratch, chunk_off, aoff_l, seg_l});
}

TIS_DEV int ceil_index(double u, int n) {   // j = ceil(u n) clamped to 1..n (fast.cpp:2257)
    int j = (int)ceil(u * n);
    if (j == n + 1) j = n;
    if (j == 0) j = 1;
    return j;
}

// ------------------------------------------------------------ k_hb
// Valence resolution: one workgroup of kResWaves wavefronts per replica.
constexpr int kResWaves = 16;
constexpr int kMaxHbGroups = 16;   // workgroups per replica in the resolution (D.hb_groups)
constexpr int kResBlock = 64 * kResWaves;
constexpr int kSeqWin = 1024;      // draw window of the sequential form (one wavefront)
constexpr int kCompWin = 192;      // draw window of each wavefront in the component form
static_assert(kResWaves * kCompWin >= kSeqWin, "window regions are shared by the two forms");
static_assert(kASiteMax <= 64 * 64, "over-subscribed sites must fit a 64 x 64 bit mask");
// LDS layout of one replica's working set (bytes, in this order):
//   E[kActMax] double | uwin[kResWaves * kCompWin] double |
//   par, excess, apack [kASiteMax] int | croot [kASiteMax] short | A1 [2 kASiteMax] short |
//   aoff[kASiteMax + 4] short | hoff[kActMax + 4] short | alist, hinc [kAIncMax] short | flag, dcls [kActMax] u8
constexpr size_t kHbLds = kActMax * 8 + kResWaves * kCompWin * 8 + 3 * kASiteMax * 4 + 3 * kASiteMax * 2 +
                          (kASiteMax + 4) * 2 + (kActMax + 4) * 2 + 2 * kAIncMax * 2 + 2 * kActMax;
static_assert(kHbLds <= 160 * 1024 - 1024, "HB working set must fit the LDS of one CU");

// FLAG_TAKEN (LDS only): switched off and taken back (k_hbres)
enum { FLAG_GATED = 0, FLAG_LIVE = 1, FLAG_OFF = 2, FLAG_TAKEN = 3 };
constexpr unsigned char kGateDode = 0x10;   // k_hbenergy's gate byte: tertiary class (E3_HB)

// HB geometry of candidate positions p[0..5] with its constants geo (r0, theta1,
// theta2, psi0, psi1, psi2): Q and the re-wrapped dihedral targets.
template <class R, class Pt>
TIS_DEV void hb_geometry_at(const Pt *pt, const double *geo, R *q_out, R *psi0) {
    const Pt &p0 = pt[0], &p1 = pt[1], &p2 = pt[2], &p3 = pt[3], &p4 = pt[4], &p5 = pt[5];
    V3<R> v14 = diff<R>(p1, p4);
    R rr = sqrt(dot(v14, v14));
    R c1, c2;
    V3<R> ga, gb;
    angle_grad<R>(diff<R>(p1, p0), v14, &c1, &ga, &gb);
    angle_grad<R>(diff<R>(p4, p3), diff<R>(p4, p1), &c2, &ga, &gb);
    R th1 = acos(unit_clamp(c1)), th2 = acos(unit_clamp(c2));
    R ps = dihedral_only<R>(diff<R>(p0, p1), diff<R>(p1, p4), diff<R>(p4, p3));
    R ps1 = dihedral_only<R>(diff<R>(p4, p1), diff<R>(p1, p0), diff<R>(p0, p2));
    R ps2 = dihedral_only<R>(diff<R>(p1, p4), diff<R>(p4, p3), diff<R>(p3, p5));
    psi0[0] = wrap_target<R>(ps, (R)geo[3]);
    psi0[1] = wrap_target<R>(ps1, (R)geo[4]);
    psi0[2] = wrap_target<R>(ps2, (R)geo[5]);
    R dr = rr - (R)geo[0];
    R q = (R)5 * dr * dr;
    q += (R)1.5 * (th1 - (R)geo[1]) * (th1 - (R)geo[1]);
    q += (R)1.5 * (th2 - (R)geo[2]) * (th2 - (R)geo[2]);
    q += (R)0.15 * (ps - psi0[0]) * (ps - psi0[0]);
    q += (R)0.15 * (ps1 - psi0[1]) * (ps1 - psi0[1]);
    q += (R)0.15 * (ps2 - psi0[2]) * (ps2 - psi0[2]);
    *q_out = q;
}
template <class R>
TIS_DEV void hb_geometry(const Dev &D, int g, R *q_out, R *psi0, V3<R> *d_unused) {
    (void)d_unused;
    const int *b = D.hb_b + 6 * g;
    using Pt = typename PtOf<R>::type;
    const Pt p[6] = {PtLoad<R>::at(D, b[0]), PtLoad<R>::at(D, b[1]), PtLoad<R>::at(D, b[2]),
                     PtLoad<R>::at(D, b[3]), PtLoad<R>::at(D, b[4]), PtLoad<R>::at(D, b[5])};
    hb_geometry_at<R>(p, D.hb_g + 6 * g, q_out, psi0);
}

// Forces of the six-bead Gaussian/Lorentzian-type term: with w the coefficient
// (-E for HB, -E/(1+Q) for stacks), each harmonic piece k (y - y0)^2 is applied
// with k scaled by w (stacking_*_force calls, fast.cpp:2043-2055, 2333-2345).
// sextet_force_parts sums the six beads' forces per bead into fb[0..5];
// sextet_forces_at adds them to F and, with vdotf, also returns the sum over the
// six beads of (force value) . (position tangent), the derivative of v . F along
// a factor that scales the whole term.
template <class R, class Pt>
TIS_DEV void sextet_force_parts(V3<R> fb[6], const Pt *p, const double *geo, R w, const R psi0[3], R kr, R kth,
                                R kpsi) {
#pragma unroll
    for (int q = 0; q < 6; ++q) fb[q] = V3<R>{(R)0, (R)0, (R)0};
    // bond (b1, b4)
    {
        V3<R> d = diff<R>(p[1], p[4]);
        R t = bond_coef<R>(d, kr * w, (R)geo[0], nullptr);
        fb[1] = fb[1] + d * (-t);
        fb[4] = fb[4] + d * t;
    }
    // angles (b0, b1, b4) and (b3, b4, b1) with the near-pi guard of stacking_valence_force
    constexpr int tri[2][3] = {{0, 1, 4}, {3, 4, 1}};
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        V3<R> ba = diff<R>(p[tri[a][1]], p[tri[a][0]]), bc = diff<R>(p[tri[a][1]], p[tri[a][2]]);
        R c;
        V3<R> g0, g2;
        angle_grad<R>(ba, bc, &c, &g0, &g2);
        if (c < (R)-0.99995) continue;
        c = unit_clamp(c);
        R s = sqrt((R)1 - c * c);
        if (!(s > (R)0)) continue;        // the reference's zero sine: no force
        R th = acos(c);
        R coef = (R)-2 * (kth * w) * ((R)geo[1 + a] - th) / s;
        V3<R> f0 = g0 * coef, f2 = g2 * coef;
        fb[tri[a][0]] = fb[tri[a][0]] + f0;
        fb[tri[a][2]] = fb[tri[a][2]] + f2;
        fb[tri[a][1]] = fb[tri[a][1]] + (f0 + f2) * (R)-1;
    }
    // dihedrals (b0,b1,b4,b3), (b4,b1,b0,b2), (b1,b4,b3,b5)
    constexpr int quad[3][4] = {{0, 1, 4, 3}, {4, 1, 0, 2}, {1, 4, 3, 5}};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        V3<R> g[4];
        R psi = dihedral_grad<R>(diff<R>(p[quad[a][0]], p[quad[a][1]]), diff<R>(p[quad[a][1]], p[quad[a][2]]),
                                 diff<R>(p[quad[a][2]], p[quad[a][3]]), g);
        R coef = (R)2 * (kpsi * w) * (psi0[a] - psi);
#pragma unroll
        for (int c = 0; c < 4; ++c) fb[quad[a][c]] = fb[quad[a][c]] + g[c] * coef;
    }
}
template <class R, class Pt, class FT>
TIS_DEV void sextet_forces_at(FT F, const int *b, const Pt *p, const double *geo, R w, const R psi0[3],
                              R kr, R kth, R kpsi, double *vdotf = nullptr) {
    // the six beads' forces are summed first, then added once each
    V3<R> fb[6];
    sextet_force_parts<R>(fb, p, geo, w, psi0, kr, kth, kpsi);
#pragma unroll
    for (int q = 0; q < 6; ++q) atomic_add_force(F, b[q], fb[q]);
    if (vdotf) {
        double sdot = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) sdot += value_dot_tangent(fb[q], p[q]);
        *vdotf = sdot;
    }
}
template <class R>
TIS_DEV void sextet_forces(const Dev &D, double4 *F, const int *b, const double *geo, R w, const R psi0[3],
                           R kr, R kth, R kpsi, double *vdotf = nullptr) {
    using Pt = typename PtOf<R>::type;
    const Pt p[6] = {PtLoad<R>::at(D, b[0]), PtLoad<R>::at(D, b[1]), PtLoad<R>::at(D, b[2]),
                     PtLoad<R>::at(D, b[3]), PtLoad<R>::at(D, b[4]), PtLoad<R>::at(D, b[5])};
    sextet_forces_at<R>(F, b, p, geo, w, psi0, kr, kth, kpsi, vdotf);
}

// Uniform source of one step's HB resolution, with the replica's counters
// hoisted into registers.  Philox: draw idx of stream `comp` is word idx % 4 of
// block (comp << 20) + idx / 4 (counter = step, block, tag).  Stream: value
// cur + idx of the caller-fed glibc stream (one stream, comp 0).
struct HbDraws {
    int rng;
    uint32_t c0, c1, k0, k1;
    const int32_t *stream;
    long long cur, lim;
    // klok, epoch and cursor: the replica's step counter, Philox epoch and stream cursor
    TIS_DEV HbDraws(const Dev &D, int r, long long klok, unsigned epoch, long long cursor) {
        rng = D.rng;
        c0 = (uint32_t)klok;
        c1 = (uint32_t)(klok >> 32);
        k0 = philox_key0(D, r);
        k1 = (uint32_t)(D.seed >> 32) + epoch;
        stream = rng ? D.stream + D.stream_base[r] : nullptr;
        cur = cursor;
        lim = rng ? D.stream_len[r] : 0;
    }
    // draws available (stream mode); draws at or past it read 0.5 and the caller
    // reports the exhausted stream
    TIS_DEV long long avail() const { return rng ? lim - cur : (1ll << 62); }
    // ubuf[i] = draw lo + i for i < n (lo and n multiples of 4); threads tid, tid + nt, ...
    TIS_DEV void fill(double *ubuf, long long lo, int n, int tid, int nt, int comp) const {
        if (rng == 0) {
            for (int q = tid; q < n / 4; q += nt) {
                const uint32_t blk = ((uint32_t)comp << 20) + (uint32_t)((lo >> 2) + q);
                const Philox4 b = philox4x32_10(c0, c1, blk, kTagHb, k0, k1);
                for (int w = 0; w < 4; ++w) ubuf[4 * q + w] = u01_closed_open(b.v[w]);
            }
        } else {
            for (int i = tid; i < n; i += nt) {
                const long long c = cur + lo + i;
                ubuf[i] = c < lim ? (double)stream[c] / kRandMax : 0.5;
            }
        }
    }
};

// u < exp(dE / T), the reference's acceptance test (fast.cpp:2300-2306), decided
// by an FP32 exponential when u lies outside +-1e-4 (relative) of it, which that
// exponential's error (<= 1e-5 relative for |dE/T| < 88) cannot cross, and by the
// FP64 expression itself otherwise.
TIS_DEV bool hb_accept(double u, double dE, double T, double invT) {
    const float e = __expf((float)(dE * invT));
    const double ed = (double)e;
    if (u < ed * (1.0 - 1e-4)) return true;
    if (u > ed * (1.0 + 1e-4)) return false;
    return u < exp(dE / T);
}

// Sites with at most this many live candidates walk with the tests on the walk's
// path only; longer walks evaluate all tests of a pass at once.
constexpr int kWalkSeq = 8;

TIS_DEV void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The reference's sequential valence resolution (fast.cpp:2239-2315) over a set
// of over-subscribed sites, run by one wavefront.  Loop: pick the j-th
// over-subscribed site of the set in site order (j = ceil(u nA)), shuffle its live
// candidates (static HB_ATOM order, multiplicity kept) with nK draws, walk them
// keeping k and moving to i when u < exp((E_i - E_k)/T), switch the pick off and
// lower the excess of every site it is listed at.
//
// The set is a bit mask over over-subscribed site indices ci (site order), word w
// held by lane w in register w0 on entry: bit b stands for index c0 + b, whose
// candidate list is apack[ci] = offset | length << 16 in alist and whose excess
// is excess[ci]; a candidate's incidences hinc hold the over-subscribed index of
// each of its sites (-1 for the others, whose excess no longer matters: it never
// rises).  The pick is a wavefront prefix count; a walk over more than kWalkSeq
// candidates evaluates all acceptance tests of a pass at once (lane t holds pair
// (i, c), c < i, triangular index t = i(i-1)/2 + c) and reads their bits, a
// shorter one evaluates the tests on its path.  Draws come from an LDS window of `win` values that the wavefront
// refills.  Sets *ovf on a site with more than 64 live candidates or an
// inconsistent table.  The switched-off candidates are recorded in offl (the
// first cap of them; *noff_out counts all).
// Pairs (i, c), c < i, of the all-at-once walk evaluated by this lane in passes
// 0..3: triangular index t = 64 p + lane = i (i - 1) / 2 + c.
TIS_DEV void walk_pairs(int pi[4], int pc[4]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int t = 64 * p + lane;
        int i = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)t)) * 0.5f);
        while (i * (i - 1) / 2 > t) --i;
        while ((i + 1) * i / 2 <= t) ++i;
        pi[p] = i;
        pc[p] = t - i * (i - 1) / 2;
    }
}

// One pass of the process at the picked site, whose candidate list is pk =
// offset | length << 16 in alist: its live candidates compacted to lanes
// 0..nK-1 in static order, the shuffle (step i swaps i with ceil(u_{1+i} nK) -
// 1, targets computed at once) and the walk (k moves to i when u_{nK+i} <
// exp((E_i - E_k)/T)).  Returns the candidate to switch off, with *nK_out, or -1
// for an inconsistent table (more than 64 entries, no live candidate).
TIS_DEV int resolve_site_pass(int pk, const short *alist, const unsigned char *flag, const double *E,
                              const double *u, double T, double invT, const int pi[4], const int pc[4],
                              int *nK_out) {
    const int lane = threadIdx.x & 63;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    const int o0 = pk & 0xffff, len = pk >> 16;
    if (len > 64) return -1;
    int bond = -1;
    bool keep = false;
    if (lane < len) { bond = alist[o0 + lane]; keep = flag[bond] == FLAG_LIVE; }
    const unsigned long long km = __ballot(keep);
    const int nK = __popcll(km);
    if (nK == 0) return -1;
    // a kept lane sends its bond to lane pos, the others to lane 63, which is
    // never read (nK == 64 only if every lane keeps)
    const int pos = __popcll(km & lt_mask);
    int bv = __builtin_amdgcn_ds_permute((keep ? pos : 63) << 2, keep ? bond : -1);
    const int Jl = lane < nK ? ceil_index(u[1 + lane], nK) - 1 : 0;
    for (int i = 0; i < nK; ++i) {
        const int tj = readlane_i(Jl, i);
        const int bi = readlane_i(bv, i), bj = readlane_i(bv, tj);
        if (lane == i) bv = bj;
        else if (lane == tj) bv = bi;
    }
    const double Ev = lane < nK ? E[bv] : 0.0;
    int k = 0;
    const int npair = nK * (nK - 1) / 2;
    if (nK > kWalkSeq && npair <= 256) {
        unsigned long long m[4] = {0, 0, 0, 0};
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            if (64 * p < npair) {
                const int i = pi[p] < nK ? pi[p] : 0, c = pi[p] < nK ? pc[p] : 0;
                const double Ei = shfl_d(Ev, i), Ec = shfl_d(Ev, c);
                const bool acc = pi[p] < nK && hb_accept(u[nK + i], Ei - Ec, T, invT);
                m[p] = __ballot(acc);
            }
        }
        for (int i = 1; i < nK; ++i) {
            const int t = i * (i - 1) / 2 + k;
            const unsigned long long mt = t < 128 ? (t < 64 ? m[0] : m[1]) : (t < 192 ? m[2] : m[3]);
            if ((mt >> (t & 63)) & 1ull) k = i;
        }
    } else {
        // a short walk: only the tests on its path, draws loaded at once
        const double ul = lane < nK ? u[nK + lane] : 0.0;
        double Ek = readlane_d(Ev, 0);
        for (int i = 1; i < nK; ++i) {
            const double Ei = readlane_d(Ev, i);
            if (hb_accept(readlane_d(ul, i), Ei - Ek, T, invT)) { k = i; Ek = Ei; }
        }
    }
    *nK_out = nK;
    return readlane_i(bv, k);
}

TIS_DEV void resolve_wave(unsigned long long w, int *excess, int c0, int nA0, const int *apack,
                          const short *alist, const short *hoff, const short *hinc, unsigned char *flag,
                          const double *E, double T, double *ubuf, int win, const HbDraws &dr, int comp,
                          short *offl, int cap, int *iters_out, int *noff_out, long long *base_out, bool *ovf) {
    const int lane = threadIdx.x & 63;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    int pi[4], pc[4];
    walk_pairs(pi, pc);
    long long base = 0, win_lo = 0;
    int iters = 0, noff = 0;
    const double invT = 1.0 / T;
    dr.fill(ubuf, 0, win, lane, 64, comp);
    wave_lds_sync();
    for (;;) {
        const int cnt = __popcll(w);
        const int incl = wave_incl_scan(cnt);
        const int nA = readlane_i(incl, 63);
        if (nA == 0) break;
        if (++iters > 1000000) { *ovf = true; break; }   // each pass removes >= 1 excess unit
        if (base + 128 > win_lo + win) {
            win_lo = base & ~3ll;
            dr.fill(ubuf, win_lo, win, lane, 64, comp);
            wave_lds_sync();
        }
        const double *u = ubuf + (base - win_lo);
        // ---- the j-th over-subscribed site of the set
        const int j = ceil_index(u[0], nA) - 1;
        const unsigned long long hit = __ballot(incl > j);
        if (hit == 0) { *ovf = true; break; }
        const int L = __builtin_ctzll(hit);
        const int jj = j - (readlane_i(incl, L) - readlane_i(cnt, L));
        const unsigned long long wl = readlane_u64(w, L);
        const unsigned long long hb = __ballot(((wl >> lane) & 1ull) && __popcll(wl & lt_mask) == jj);
        if (hb == 0) { *ovf = true; break; }
        const int ci = c0 + 64 * L + __builtin_ctzll(hb);
        if (ci >= nA0) { *ovf = true; break; }
        int nK = 0;
        const int off = resolve_site_pass(apack[ci], alist, flag, E, u, T, invT, pi, pc, &nK);
        if (off < 0) { *ovf = true; break; }
        base += 2 * nK;
        // ---- switch it off (recorded while the record has room); sites whose
        // excess reaches 0 leave the mask
        if (lane == 0) {
            flag[off] = FLAG_OFF;
            if (noff < cap) offl[noff] = (short)off;
        }
        ++noff;
        const int h0 = hoff[off], hn = hoff[off + 1] - h0;
        if (hn > 64) { *ovf = true; break; }
        int b = -1;
        bool cleared = false;
        if (lane < hn) {
            const int ci2 = hinc[h0 + lane];   // over-subscribed index of the site, or -1
            if (ci2 >= 0) {
                cleared = atomicSub(&excess[ci2], 1) == 1;
                b = ci2 - c0;
            }
        }
        unsigned long long cm = __ballot(cleared && b >= 0 && b < 64 * 64);
        while (cm) {
            const int l = __builtin_ctzll(cm);
            cm &= cm - 1ull;
            const int bb = readlane_i(b, l);
            if (lane == (bb >> 6)) w &= ~(1ull << (bb & 63));
        }
        wave_lds_sync();
    }
    if (base > dr.avail()) *ovf = true;
    *iters_out = iters;
    *noff_out = noff;
    *base_out = base;
}

// resolve_wave for a component of at most 64 over-subscribed sites, with their
// state in registers: lane t < nm holds the component's t-th site in index order
// (mci), its excess (ex) and its candidate list (pk).  The same draws, picks and
// decisions as resolve_wave on the component's mask: the j-th over-subscribed
// site in index order is the j-th lane whose excess is positive, and every
// over-subscribed site of a switched-off candidate is a member (a component
// joins all the sites its live candidates touch), so the candidate's incidences
// lower lane excesses only (once per incidence, multiplicity kept).
TIS_DEV void resolve_small(int nm, int mci, int ex, int pk, const short *alist, const short *hoff,
                           const short *hinc, unsigned char *flag, const double *E, double T, double *ubuf,
                           int win, const HbDraws &dr, int comp, short *offl, int cap, int *iters_out,
                           int *noff_out, long long *base_out, bool *ovf) {
    const int lane = threadIdx.x & 63;
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    int pi[4], pc[4];
    walk_pairs(pi, pc);
    long long base = 0, win_lo = 0;
    int iters = 0, noff = 0;
    const double invT = 1.0 / T;
    int exv = lane < nm ? ex : 0;
    dr.fill(ubuf, 0, win, lane, 64, comp);
    wave_lds_sync();
    for (;;) {
        const unsigned long long om = __ballot(exv > 0);
        const int nA = __popcll(om);
        if (nA == 0) break;
        if (++iters > 1000000) { *ovf = true; break; }   // each pass removes >= 1 excess unit
        if (base + 128 > win_lo + win) {
            win_lo = base & ~3ll;
            dr.fill(ubuf, win_lo, win, lane, 64, comp);
            wave_lds_sync();
        }
        const double *u = ubuf + (base - win_lo);
        // ---- the j-th over-subscribed site: the lane holding it
        const int j = ceil_index(u[0], nA) - 1;
        const unsigned long long hit = __ballot(((om >> lane) & 1ull) && __popcll(om & lt_mask) == j);
        if (hit == 0) { *ovf = true; break; }
        const int L = __builtin_ctzll(hit);
        int nK = 0;
        const int off = resolve_site_pass(readlane_i(pk, L), alist, flag, E, u, T, invT, pi, pc, &nK);
        if (off < 0) { *ovf = true; break; }
        base += 2 * nK;
        // ---- switch it off (recorded while the record has room) and lower the
        // excess of its sites
        if (lane == 0) {
            flag[off] = FLAG_OFF;
            if (noff < cap) offl[noff] = (short)off;
        }
        ++noff;
        const int h0 = hoff[off], hn = hoff[off + 1] - h0;
        if (hn > 64) { *ovf = true; break; }
        const int ci2 = lane < hn ? hinc[h0 + lane] : -1;   // over-subscribed index, or -1
        for (int q = 0; q < hn; ++q) {
            const int c = readlane_i(ci2, q);
            if (lane < nm && mci == c) exv -= 1;
        }
        wave_lds_sync();
    }
    if (base > dr.avail()) *ovf = true;
    *iters_out = iters;
    *noff_out = noff;
    *base_out = base;
}

// Union-find root with path halving; par[] lives in LDS and is hooked
// concurrently (roots only ever move to smaller indices).
TIS_DEV int uf_find(volatile int *par, int x) {
    int p = par[x];
    while (p != x) {
        const int g = par[p];
        if (g != p) par[x] = g;
        x = p;
        p = par[x];
    }
    return x;
}
TIS_DEV void uf_union(volatile int *par, int a, int b) {
    int ra = uf_find(par, a), rb = uf_find(par, b);
    while (ra != rb) {
        if (ra < rb) { const int t = ra; ra = rb; rb = t; }   // hook the larger root under the smaller
        const int old = atomicCAS((int *)&par[ra], ra, rb);
        if (old == ra) break;
        ra = uf_find(par, old);
        rb = uf_find(par, rb);
    }
}

// Sums of the resolution groups of each replica (one thread per replica): the
// HB energies and the serial depth.  Steps do the same in their close.
TIS_DEV void hb_reduce_parts(const Dev &D, int r) {
    RepState &S = D.rs[r];
    double e2 = 0.0, e3 = 0.0;
    int m = 0;
    for (int q = 0; q < D.hb_groups; ++q) {
        e2 += D.hb_e_part[2 * (kMaxHbGroups * r + q)];
        e3 += D.hb_e_part[2 * (kMaxHbGroups * r + q) + 1];
        m = max(m, D.hb_iter_part[kMaxHbGroups * r + q]);
    }
    S.e[EI_E2_HB] = e2;
    S.e[EI_E3_HB] = e3;
    S.hb_iters = m;
}
__global__ void __launch_bounds__(kBlock) k_hbsum(Dev D) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < D.n_rep) hb_reduce_parts(D, r);
}

// Gate, energy and forces of every live HB candidate (hydrogen_bonds_interactions,
// fast.cpp:2180-2237, forces :2333-2345), one thread per candidate, before the
// valence resolution decides which are kept: E = E_hb exp(-Q) when |r - r0| <= 2 A
// for the (k11, k21) pair (gate flag FLAG_LIVE), else 0 (FLAG_GATED); inside the
// gate the six bead forces, which k_hbres adds to the step's multi-body forces
// (and takes back from the candidates it switches off).  The outcome (hb_out,
// act_flag) is the gate until the resolution changes it.  Computed in precision R
// once per candidate, for every resolution group.  Grid (kActMax / kBlock,
// replicas); outputs at r * kActMax + live index, the gate byte with kGateDode
// set for the tertiary class (E3_HB).
// Candidate a (live index) of replica r; n_act and tab: the replica's current
// table counts and buffer.
template <class R, class L>
TIS_DEV void hbenergy_one(const Dev &D, int r, int a, int n_act, int tab) {
    const RepParams &P = D.rp[r];
    if (a >= n_act) return;
    const HbTab T = hb_tab(D, tab);
    const int4 ra = L::ld(T.act_rec + 2 * (P.hb0 + a)), rb = L::ld(T.act_rec + 2 * (P.hb0 + a) + 1);
    const int g = ra.x;
    const double4 pt[6] = {L::ld(D.pos + ra.y), L::ld(D.pos + ra.z), L::ld(D.pos + ra.w),
                           L::ld(D.pos + rb.x), L::ld(D.pos + rb.y), L::ld(D.pos + rb.z)};
    const double *geo = D.hb_g + 6 * g;
    const double hE = D.hb_E[g];
    const unsigned char cls = D.hb_dode[g] ? kGateDode : 0;
    const size_t o = (size_t)kActMax * r + a;
    const V3<R> v = diff<R>(pt[1], pt[4]);
    const R rr = sqrt(dot(v, v));
    double e = 0.0;
    unsigned char f = FLAG_GATED;
    if (!(fabs(rr - (R)geo[0]) > (R)2)) {
        R q, psi0[3];
        hb_geometry_at<R>(pt, geo, &q, psi0);
        e = (double)((R)hE * exp(-q));
        f = FLAG_LIVE;
        V3<R> fb[6];
        sextet_force_parts<R>(fb, pt, geo, (R)(-e), psi0, (R)5, (R)1.5, (R)0.15);
        R *out = (R *)D.hb_frc + 18 * o;
#pragma unroll
        for (int q6 = 0; q6 < 6; ++q6) {
            out[3 * q6] = fb[q6].x;
            out[3 * q6 + 1] = fb[q6].y;
            out[3 * q6 + 2] = fb[q6].z;
        }
    }
    D.hb_en[o] = e;
    D.hb_gate[o] = f | cls;
    D.hb_out[g] = (signed char)f;
    D.act_flag[P.hb0 + a] = f;
}

// The six bead forces k_hbenergy kept for live candidate a (a < 0: none), added
// to the step's multi-body forces, or, for a candidate the resolution switched
// off (off), subtracted again, with its outcome set to FLAG_OFF.
template <class R, class L>
struct HbForces {
    const Dev *D;
    const int4 *rec;             // T.act_rec + 2 hb0: records by live index
    const R *frc;                // the replica's kept forces, 18 per live index
    unsigned char *act_flag;     // D.act_flag + hb0
    TIS_DEV void operator()(int a, bool off) const {
        if (a < 0) return;
        const int4 ra = L::ld(rec + 2 * a), rb = L::ld(rec + 2 * a + 1);
        const R *fr = frc + 18 * (size_t)a;
        const R sg = off ? (R)-1 : (R)1;
        V3<R> fb[6];
#pragma unroll
        for (int q = 0; q < 6; ++q)
            fb[q] = V3<R>{sg * L::ld(fr + 3 * q), sg * L::ld(fr + 3 * q + 1), sg * L::ld(fr + 3 * q + 2)};
        if (off) {
            D->hb_out[ra.x] = (signed char)FLAG_OFF;
            act_flag[a] = FLAG_OFF;
        }
        const int bd[6] = {ra.y, ra.z, ra.w, rb.x, rb.y, rb.z};
#pragma unroll
        for (int q = 0; q < 6; ++q) atomic_add_force(D->f_mb, bd[q], fb[q]);
    }
};
template <class R>
__global__ void __launch_bounds__(kBlock) k_hbenergy(Dev D) {
    const int r = blockIdx.y;
    const RepState &S = D.rs[r];
    const int n_act = (S.overflow & 2) ? 0 : S.n_act;
    hbenergy_one<R, LdPlain>(D, r, blockIdx.x * blockDim.x + threadIdx.x, n_act, S.tab);
}

// Valence resolution of the live HB candidates (hydrogen_bonds_interactions,
// fast.cpp:2239-2315) after k_hbenergy's gates, energies and forces; switches
// off candidates until no site exceeds its valence, takes back what k_hbenergy
// added for them and sums the kept candidates' energies.  mode bit 0: a step
// (advance the stream cursor past the HB draws and build the noise table).
//
// Reference stream: the exact sequential process on wavefront 0.  Philox draws:
// the over-subscribed sites split into the connected components of the graph
// whose edges are the live candidates they share (union-find), and each
// component runs its own copy of the process on one wavefront, with its own
// Philox stream.  An iteration of the reference's process picks a uniformly
// random over-subscribed site and its update touches only that site's
// component, so, conditioned on landing in component C, it is a uniform pick
// among C's over-subscribed sites: the components' processes give the same joint
// distribution of switched-off candidates as the single sequential loop, with
// the serial depth of the largest wavefront's share instead of the total.
//
// Grid (groups, replicas).  With Philox draws the components of a replica are
// shared out over the groups (component k to group k mod groups, each group's
// wavefronts taking its components in turn); every group builds the tables and
// components itself, and the wavefront that ran a component takes back the
// candidates it switched off as soon as it is done.  Each component draws from
// its own Philox stream, so the outcome does not depend on the number of groups.
// The reference-stream process runs in group 0.  Energy sums: group 0 adds
// every live candidate's energy, each group subtracts those it switched off.
// Group grp of ngrp of replica r's resolution, a workgroup of kResBlock
// threads with kHbLds bytes of LDS at lds.
template <class R, class L>
TIS_DEV void hbres_group(const Dev &D, int mode, int r, int grp, int ngrp, unsigned char *lds) {
    RepState &S = D.rs[r];
    const RepParams &P = D.rp[r];
    double *E = (double *)lds;
    double *uwin = E + kActMax;
    int *par = (int *)(uwin + kResWaves * kCompWin);
    int *excess = par + kASiteMax;
    int *apack = excess + kASiteMax;
    int *cnt = apack;   // live incidences per active site, until the ranking writes apack
    short *croot = (short *)(apack + kASiteMax);
    short *A1 = croot + kASiteMax;
    short *offl = A1;   // 2 kASiteMax entries: each wavefront's switched-off candidates, once A1 is read
    short *aoff = A1 + 2 * kASiteMax;
    short *hoff = aoff + kASiteMax + 4;
    short *alist = hoff + kActMax + 4;
    short *hinc = alist + kAIncMax;
    unsigned char *flag = (unsigned char *)(hinc + kAIncMax);
    unsigned char *dcls = flag + kActMax;   // kGateDode of each live candidate
    __shared__ int scratch[20];
    __shared__ int sh_iters;
    __shared__ long long sh_draws;
    __shared__ int sh_err, sh_next, sh_fnext;
    const int n_act = L::ld(&S.n_act), n_as = L::ld(&S.n_asite), n_ainc = L::ld(&S.n_ainc);
    const HbTab T = hb_tab(D, L::ld(&S.tab));
    // phase clock (tis_trace_hb): thread 0 stamps the end of each phase
    unsigned long long *trc = nullptr;
    if (D.hb_trace && threadIdx.x == 0) {
        const long long slot = L::ld(&S.klok) % kHbTraceSteps;
        trc = D.hb_trace + (((size_t)slot * D.n_rep + r) * kMaxHbGroups + grp) * kHbTraceWords;
        trc[0] = __builtin_amdgcn_s_memrealtime();
    }
    auto stamp = [&](int k) {
        if (trc) trc[k] = __builtin_amdgcn_s_memrealtime();
    };
    if (threadIdx.x == 0) {   // this group's parts (summed by k_hbsum or the step's close)
        D.hb_iter_part[kMaxHbGroups * r + grp] = 0;
        D.hb_e_part[2 * (kMaxHbGroups * r + grp)] = 0.0;
        D.hb_e_part[2 * (kMaxHbGroups * r + grp) + 1] = 0.0;
    }
    if (L::ld(&S.overflow) & 2) return;
    if (D.rng == 1 && grp > 0) return;
    if (threadIdx.x == 0) { sh_err = 0; sh_draws = 0; sh_iters = 0; sh_next = kResWaves; sh_fnext = 0; }
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6;
    // ---- the rebuild tables with the active sites' valences, and the candidate
    // energies and gates of k_hbenergy: one level of loads, all issued before any
    // is used.  The live incidences counted from the sites (n_ainc) are those
    // counted from the candidates.
    if (n_ainc > kAIncMax) {                      // inconsistent rebuild tables
        if (tid == 0) atomicOr(&S.overflow, 2);
        return;
    }
    constexpr int KA = (kASiteMax + kResBlock) / kResBlock;   // slots per thread: sites (+1)
    constexpr int KC = (kActMax + kResBlock) / kResBlock;     // live candidates (+1)
    constexpr int KI = (kAIncMax + kResBlock - 1) / kResBlock;  // incidences
    int t_aoff[KA], t_val[KA], t_hoff[KC], t_alist[KI], t_hinc[KI];
    double t_e[KC];
    unsigned char t_f[KC];
#pragma unroll
    for (int k = 0; k < KA; ++k) {
        const int i = tid + k * kResBlock;
        t_aoff[k] = i <= n_as ? L::ld(T.asite_off + P.site0 + r + i) : 0;
        t_val[k] = i < n_as ? L::ld(T.asite_val + P.site0 + i) : 0;
    }
    // energies and gates of the live candidates (k_hbenergy)
    const double *hb_en = D.hb_en + (size_t)kActMax * r;
    const unsigned char *hb_gate = D.hb_gate + (size_t)kActMax * r;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        const int i = tid + k * kResBlock;
        t_hoff[k] = i <= n_act ? L::ld(T.ahb_inc_off + P.hb0 + r + i) : 0;
        t_e[k] = i < n_act ? L::ld(hb_en + i) : 0.0;
        t_f[k] = i < n_act ? L::ld(hb_gate + i) : (unsigned char)FLAG_GATED;
    }
#pragma unroll
    for (int k = 0; k < KI; ++k) {
        const int i = tid + k * kResBlock;
        t_alist[k] = i < n_ainc ? L::ld(T.asite_list + P.inc0 + i) : 0;
        t_hinc[k] = i < n_ainc ? L::ld(T.ahb_inc + P.inc0 + i) : 0;
    }
#pragma unroll
    for (int k = 0; k < KA; ++k) {
        const int i = tid + k * kResBlock;
        if (i <= n_as) aoff[i] = (short)t_aoff[k];
        if (i < n_as) { excess[i] = t_val[k]; cnt[i] = 0; }   // the valence, until the excess replaces it
    }
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        const int i = tid + k * kResBlock;
        if (i <= n_act) hoff[i] = (short)t_hoff[k];
        if (i < n_act) {
            E[i] = t_e[k];
            flag[i] = t_f[k] & 0x0f;
            dcls[i] = t_f[k] & kGateDode;
        }
    }
#pragma unroll
    for (int k = 0; k < KI; ++k) {
        const int i = tid + k * kResBlock;
        if (i < n_ainc) {
            alist[i] = (short)t_alist[k];
            hinc[i] = (short)t_hinc[k];
        }
    }
    __syncthreads();
    stamp(1);
    // ---- live incidences per active site: each live candidate's thread adds
    // one at every site it lists (with multiplicity, as the sites' lists hold it)
    int t_hend[KC];
    bool t_live[KC];
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        const int a = tid + k * kResBlock;
        t_hend[k] = a < n_act ? hoff[a + 1] : 0;
        t_live[k] = a < n_act && flag[a] == FLAG_LIVE;
    }
#pragma unroll
    for (int k = 0; k < KC; ++k) {
        if (t_live[k]) {
            const int q0 = t_hoff[k], nq = t_hend[k] - q0;
            int st[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) st[j] = j < nq ? hinc[q0 + j] : 0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (j < nq) atomicAdd(&cnt[st[j]], 1);
            for (int q = q0 + 4; q < q0 + nq; ++q) atomicAdd(&cnt[hinc[q]], 1);
        }
    }
    __syncthreads();
    // ---- excess = live incidences - valence per active site; the
    // over-subscribed ones in site order (index ci: candidate list apack[ci]) and
    // the inverse map A1.  Each thread takes a contiguous range of sites, so one
    // scan ranks them in site order.
    constexpr int KS = (kASiteMax + kResBlock - 1) / kResBlock;   // sites per thread
    int nA0;
    {
        const int per = (n_as + nt - 1) / nt;
        const int s_lo = min(n_as, tid * per);
        int exs[KS], mine = 0;
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            const int s = s_lo + k;
            exs[k] = 0;
            if (k < per && s < n_as) {
                exs[k] = cnt[s] - excess[s];
                mine += exs[k] > 0;
            }
        }
        // excess re-indexed by over-subscribed index (rank <= s); every count and
        // valence was read before the scan's barriers
        int rank = block_excl_sum(mine, scratch, &nA0);
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            const int s = s_lo + k;
            if (k < per && s < n_as) {
                const int ex = exs[k];
                A1[s] = (short)(ex > 0 ? rank : -1);
                if (ex > 0) {
                    apack[rank] = (int)aoff[s] | ((int)(aoff[s + 1] - aoff[s]) << 16);
                    excess[rank] = ex;
                    par[rank] = rank;   // union-find roots of the component form
                    ++rank;
                }
            }
        }
    }
    __syncthreads();
    stamp(2);
    int ncomp = 0;
    const HbDraws dr(D, r, L::ld(&S.klok), L::ld(&S.rng_epoch), L::ld(&S.stream_cursor));
    const HbForces<R, L> frc{&D, T.act_rec + 2 * P.hb0, (const R *)D.hb_frc + 18 * (size_t)kActMax * r,
                             D.act_flag + P.hb0};
    // The forces of every live candidate, added by the wavefronts that have no
    // (more) components to run, while the others resolve theirs: this group's
    // share of the live indices, in chunks of 64 handed out by sh_fnext
    auto add_live = [&]() {
        const bool all = D.rng == 1;
        const int lo = all ? 0 : (int)((long long)n_act * grp / ngrp);
        const int hi = all ? n_act : (int)((long long)n_act * (grp + 1) / ngrp);
        for (;;) {
            int c = 0;
            if (lane == 0) c = atomicAdd(&sh_fnext, 1);
            c = __shfl(c, 0, 64);
            const int a = lo + 64 * c + lane;
            if (lo + 64 * c >= hi) break;
            // the gate never changes here (the resolution moves live to off)
            frc(a < hi && flag[a] != FLAG_GATED ? a : -1, false);
        }
    };
    // A wavefront takes back the candidates it switched off: the n it recorded
    // (up to cap), then, if the record filled up, the rest, found by their outcome
    // and their component (root; -1: every over-subscribed site)
    auto take_wave = [&](const short *rec, int n, int cap, int root) {
        const int nr = min(n, cap);
        for (int i0 = 0; i0 < nr; i0 += 64) {
            const int a = i0 + lane < nr ? rec[i0 + lane] : -1;
            frc(a, true);
            if (a >= 0) flag[a] = FLAG_TAKEN;
        }
        if (n > cap) {
            for (int a0 = 0; a0 < n_act; a0 += 64) {
                const int a = a0 + lane;
                bool mine = a < n_act && flag[a] == FLAG_OFF;
                if (mine && root >= 0) {
                    int ci = -1;
                    for (int q = hoff[a]; q < hoff[a + 1] && ci < 0; ++q) ci = hinc[q];
                    mine = ci >= 0 && par[ci] == root;
                }
                frc(mine ? a : -1, true);
                if (mine) flag[a] = FLAG_TAKEN;
            }
        }
        wave_lds_sync();
    };
    // ---- each live candidate's thread turns its incidences into over-subscribed
    // indices (-1: a site that is not) and, with Philox draws, joins the
    // over-subscribed sites it touches (union-find over A0 indices; roots set with
    // the excess).  Only live candidates are ever switched off, so only their
    // incidences are read later.
    {
        volatile int *vpar = par;
#pragma unroll
        for (int k = 0; k < KC; ++k) {
            if (t_live[k]) {
                const int q0 = t_hoff[k], nq = t_hend[k] - q0;
                int st[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) st[j] = j < nq ? hinc[q0 + j] : 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) st[j] = j < nq ? A1[st[j]] : -1;
                for (int q = q0 + 4; q < q0 + nq; ++q) hinc[q] = A1[hinc[q]];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (j < nq) hinc[q0 + j] = (short)st[j];
                if (D.rng == 0) {
                    int first = -1;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (j >= nq || st[j] < 0) continue;
                        if (first < 0) first = st[j];
                        else uf_union(vpar, first, st[j]);
                    }
                    for (int q = q0 + 4; q < q0 + nq; ++q) {
                        const int ci = hinc[q];
                        if (ci < 0) continue;
                        if (first < 0) first = ci;
                        else uf_union(vpar, first, ci);
                    }
                }
            }
        }
    }
    __syncthreads();
    stamp(3);
    if (D.rng == 1) {
        stamp(4);
        // reference stream: the exact sequential process on wavefront 0, over all
        // over-subscribed sites
        if (wv == 0) {
            const int c = 64 * lane;
            const unsigned long long w0 = c + 64 <= nA0 ? ~0ull : (c < nA0 ? (1ull << (nA0 - c)) - 1ull : 0ull);
            int iters = 0, noff = 0;
            long long base = 0;
            bool ovf = false;
            resolve_wave(w0, excess, 0, nA0, apack, alist, hoff, hinc, flag, E, P.T, uwin, kSeqWin, dr, 0, offl,
                         2 * kASiteMax, &iters, &noff, &base, &ovf);
            take_wave(offl, noff, 2 * kASiteMax, -1);
            if (lane == 0) {
                sh_iters = iters;
                sh_draws = base;
                if (ovf) sh_err = 1;
            }
        }
        add_live();
    } else {
        volatile int *vpar = par;
        // roots (smallest index of each component) in order, each thread over a
        // contiguous range of indices; every index points at its root afterwards
        // (finds only shorten paths, so they may run beside each other)
        {
            const int per = (nA0 + nt - 1) / nt;
            const int c_lo = min(nA0, tid * per);
            int rts[KS], mine = 0;
#pragma unroll
            for (int k = 0; k < KS; ++k) {
                const int ci = c_lo + k;
                rts[k] = -1;
                if (k < per && ci < nA0) {
                    rts[k] = uf_find(vpar, ci);
                    par[ci] = rts[k];
                    mine += rts[k] == ci;
                }
            }
            int rank = block_excl_sum(mine, scratch, &ncomp);
#pragma unroll
            for (int k = 0; k < KS; ++k) {
                const int ci = c_lo + k;
                if (k < per && ci < nA0 && rts[k] == ci) {
                    croot[rank] = (short)ci;
                    ++rank;
                }
            }
        }
        __syncthreads();
        stamp(4);
        // ---- each wavefront takes the next component in root order when it is done
        double *ub = uwin + wv * kCompWin;
        constexpr int kOffCap = 2 * kASiteMax / kResWaves;
        short *offw = offl + wv * kOffCap;
        int wave_iters = 0;
        bool ovf = false;
        for (int j = wv;;) {
            const int k = grp + ngrp * j;
            if (k >= ncomp) {
                add_live();
                break;
            }
            const int root = croot[k], cfirst = root >> 6, c0 = cfirst << 6;
            // member mask (word w from the ballot over indices c0 + 64 w .. + 63)
            // and, up to 64 members, the members in index order in lanes 0, 1, ...
            unsigned long long m = 0;
            int nm = 0, mci = -1;
            const int clast = (nA0 - 1) >> 6;
            for (int c = cfirst; c <= clast; ++c) {
                const int ci = 64 * c + lane;
                unsigned long long bm = __ballot(ci < nA0 && par[ci] == root);
                if (lane == c - cfirst) m = bm;
                while (bm) {
                    const int b = __builtin_ctzll(bm);
                    bm &= bm - 1ull;
                    if (lane == nm) mci = 64 * c + b;
                    ++nm;
                }
            }
            int it = 0, noff = 0;
            long long base = 0;
            if (nm <= 64) {
                const int exm = lane < nm ? excess[mci] : 0, pkm = lane < nm ? apack[mci] : 0;
                resolve_small(nm, mci, exm, pkm, alist, hoff, hinc, flag, E, P.T, ub, kCompWin, dr, k, offw,
                              kOffCap, &it, &noff, &base, &ovf);
            } else {
                resolve_wave(m, excess, c0, nA0, apack, alist, hoff, hinc, flag, E, P.T, ub, kCompWin, dr, k, offw,
                             kOffCap, &it, &noff, &base, &ovf);
            }
            take_wave(offw, noff, kOffCap, root);
            wave_iters += it;
            if (ovf) break;
            int nj = 0;
            if (lane == 0) nj = atomicAdd(&sh_next, 1);
            j = __shfl(nj, 0, 64);
        }
        if (lane == 0) {
            atomicMax(&sh_iters, wave_iters);
            if (ovf) sh_err = 1;
        }
    }
    __syncthreads();
    stamp(5);
    if (trc) {
        trc[8] = (unsigned long long)sh_iters;
        trc[9] = (unsigned long long)nA0 | (unsigned long long)ncomp << 16 | (unsigned long long)n_act << 32;
    }
    // ---- energy sums (E2_HB, E3_HB): group 0 adds every live candidate's
    // energy, every group subtracts those it switched off (their forces were
    // taken back by the wavefront that ran their component)
    {
        double e2 = 0.0, e3 = 0.0;
        for (int a = tid; a < n_act; a += nt) {
            const unsigned char f = flag[a];
            if (f == FLAG_GATED) continue;
            const double e = E[a];
            double v = grp == 0 ? e : 0.0;
            if (f >= FLAG_OFF) v -= e;
            if (dcls[a]) e3 += v; else e2 += v;
        }
        __shared__ double red[kResWaves];   // one per wavefront of the block
        const double s2 = block_sum_double(e2, red);
        const double s3 = block_sum_double(e3, red);
        if (tid == 0) {
            D.hb_e_part[2 * (kMaxHbGroups * r + grp)] = s2;
            D.hb_e_part[2 * (kMaxHbGroups * r + grp) + 1] = s3;
        }
    }
    stamp(6);
    // first resolution on a table switched by the step's close: candidates outside it
    // (inactive in the rebuild that made it) report -1; k_hbenergy wrote the
    // listed ones
    if (L::ld(&S.tab_fresh)) {
        const int f0 = D.rng == 1 ? 0 : grp, fs = D.rng == 1 ? 1 : ngrp;
        for (int k = f0 * nt + tid; k < P.nhb; k += fs * nt)
            if (L::ld(D.hb_status + P.hb0 + k) < 0) D.hb_out[P.hb0 + k] = -1;
    }
    if (tid == 0) {
        D.hb_iter_part[kMaxHbGroups * r + grp] = sh_iters;
        if (D.rng == 1) S.hb_draws = sh_draws;
        if (sh_err) atomicOr(&S.overflow, D.rng ? 4 : 16);
    }
    stamp(7);
    // ---- reference-stream noise table: 1500 Box-Muller pairs after the HB draws
    // (full_forces, fast.cpp:2598-2611).  r1 == 0 is redrawn in the reference; when
    // no r1 slot of this step's block holds a 0 (all but ~1e-6 of steps) the pairs
    // are independent and are formed in parallel, otherwise one thread walks them.
    if ((mode & 1) && D.rng == 1) {
        __shared__ int sh_flags;
        if (threadIdx.x == 0) sh_flags = 0;
        __syncthreads();
        const long long c0 = L::ld(&S.stream_cursor) + sh_draws;
        const long long lim = D.stream_len[r];
        const int32_t *st = D.stream + D.stream_base[r];
        double *mw = D.maxwell + 3000 * (size_t)r;
        for (int q = threadIdx.x; q < 1500; q += blockDim.x) {
            const long long c = c0 + 2 * q;
            if (c < 0 || c + 1 >= lim) atomicOr(&sh_flags, 2);
            else if (st[c] == 0) atomicOr(&sh_flags, 1);
        }
        __syncthreads();
        if (sh_flags == 0) {
            for (int q = threadIdx.x; q < 1500; q += blockDim.x) {
                const long long c = c0 + 2 * q;
                const double r1 = (double)st[c] / kRandMax, r2 = (double)st[c + 1] / kRandMax;
                const double a = sqrt(-2.0 * log(r1)), b = 2.0 * kPi * r2;
                mw[2 * q] = a * cos(b);
                mw[2 * q + 1] = a * sin(b);
            }
            __syncthreads();
            if (threadIdx.x == 0) S.stream_cursor = c0 + 3000;
        } else if (threadIdx.x == 0) {
            long long c = c0;
            bool bad = false;
            for (int i = 0; i < 3000; i += 2) {
                double r1;
                do {
                    if (c < 0 || c >= lim) { bad = true; r1 = 0.5; break; }
                    r1 = (double)st[c++] / kRandMax;
                } while (r1 == 0);
                double r2 = 0.5;
                if (c >= 0 && c < lim) r2 = (double)st[c++] / kRandMax; else bad = true;
                const double a = sqrt(-2.0 * log(r1)), b = 2.0 * kPi * r2;
                mw[i] = a * cos(b);
                mw[i + 1] = a * sin(b);
            }
            S.stream_cursor = c;
            if (bad) atomicOr(&S.overflow, 4);
        }
    } else if (threadIdx.x == 0 && D.rng == 1) {
        S.stream_cursor += sh_draws;   // evaluation consumes the HB draws only
    }
}
template <class R>
__global__ void __launch_bounds__(kResBlock) k_hbres(Dev D, int mode) {
    extern __shared__ unsigned char lds[];
    hbres_group<R, LdPlain>(D, mode, blockIdx.y, blockIdx.x, gridDim.x, lds);
}

// Tertiary stack g (stacking_interactions, fast.cpp:1911-2141): E / (1 + Q6)
// with Q6 the weighted squared deviations of one distance, two angles and three
// dihedrals; returns the energy and adds the forces (see sextet_forces for vdotf).
template <class R, class L = LdPlain, class FT>
TIS_DEV R tertiary_stack(const Dev &D, FT F, int g, double *vdotf) {
    using Pt = typename PtOf<R>::type;
    const int *b = D.st3_b + 6 * g;
    const double *geo = D.st3_g + 6 * g;
    const Pt p0 = PtLoad<R, L>::at(D, b[0]), p1 = PtLoad<R, L>::at(D, b[1]), p2 = PtLoad<R, L>::at(D, b[2]);
    const Pt p3 = PtLoad<R, L>::at(D, b[3]), p4 = PtLoad<R, L>::at(D, b[4]), p5 = PtLoad<R, L>::at(D, b[5]);
    V3<R> v14 = diff<R>(p1, p4);
    R dr = sqrt(dot(v14, v14)) - (R)geo[0];
    R c1, c2;
    V3<R> ga, gb;
    angle_grad<R>(diff<R>(p1, p0), v14, &c1, &ga, &gb);
    angle_grad<R>(diff<R>(p4, p3), diff<R>(p4, p1), &c2, &ga, &gb);
    R th1 = acos(unit_clamp(c1)), th2 = acos(unit_clamp(c2));
    R ps = dihedral_only<R>(diff<R>(p0, p1), diff<R>(p1, p4), diff<R>(p4, p3));
    R ps1 = dihedral_only<R>(diff<R>(p4, p1), diff<R>(p1, p0), diff<R>(p0, p2));
    R ps2 = dihedral_only<R>(diff<R>(p1, p4), diff<R>(p4, p3), diff<R>(p3, p5));
    R psi0[3] = {wrap_target<R>(ps, (R)geo[3]), wrap_target<R>(ps1, (R)geo[4]), wrap_target<R>(ps2, (R)geo[5])};
    R q = (R)1 + (R)5 * dr * dr;
    q += (R)1.5 * (th1 - (R)geo[1]) * (th1 - (R)geo[1]);
    q += (R)1.5 * (th2 - (R)geo[2]) * (th2 - (R)geo[2]);
    q += (R)0.15 * (ps - psi0[0]) * (ps - psi0[0]);
    q += (R)0.15 * (ps1 - psi0[1]) * (ps1 - psi0[1]);
    q += (R)0.15 * (ps2 - psi0[2]) * (ps2 - psi0[2]);
    const R e = (R)D.st3_E[g] / q;
    const Pt pt[6] = {p0, p1, p2, p3, p4, p5};
    sextet_forces_at<R>(F, b, pt, geo, -e / q, psi0, (R)5, (R)1.5, (R)0.15, vdotf);
    return e;
}

// Secondary stack g (stacking_interactions, fast.cpp:2060-2141): E / (1 + Q3)
// over the B-B distance and two backbone dihedrals; returns the energy and adds
// the forces (vdotf as in sextet_forces).
template <class R, class L = LdPlain, class FT>
TIS_DEV R secondary_stack(const Dev &D, FT F, int g, double *vdotf) {
    using Pt = typename PtOf<R>::type;
    const int *b = D.st2_b + 8 * g;   // P_n S_n B_n P_n+1 S_n+1 B_n+1 P_n+2
    const double2 rE = D.st2_rE[g];
    const Pt q0 = PtLoad<R, L>::at(D, b[0]), q1 = PtLoad<R, L>::at(D, b[1]), q2 = PtLoad<R, L>::at(D, b[2]);
    const Pt q3 = PtLoad<R, L>::at(D, b[3]), q4 = PtLoad<R, L>::at(D, b[4]), q5 = PtLoad<R, L>::at(D, b[5]);
    const Pt q6 = PtLoad<R, L>::at(D, b[6]);
    V3<R> v25 = diff<R>(q2, q5);
    R rr = sqrt(dot(v25, v25));
    V3<R> g1[4], g2[4];
    R ps1 = dihedral_grad<R>(diff<R>(q0, q1), diff<R>(q1, q3), diff<R>(q3, q4), g1);
    R ps2 = dihedral_grad<R>(diff<R>(q6, q4), diff<R>(q4, q3), diff<R>(q3, q1), g2);
    R t1 = wrap_target<R>(ps1, (R)-2.58684), t2 = wrap_target<R>(ps2, (R)3.07135);
    R q = (R)1 + (R)1.4 * (rr - (R)rE.x) * (rr - (R)rE.x);
    q += (R)4 * (ps1 - t1) * (ps1 - t1);
    q += (R)4 * (ps2 - t2) * (ps2 - t2);
    const R e = (R)rE.y / q;
    const R w = -e / q;
    const R tb = (R)2 * ((R)1.4 * w) * ((R)rE.x - rr) / rr;
    double sdot = 0.0;
    const V3<R> f2 = v25 * (-tb), f5 = v25 * tb;
    atomic_add_force(F, b[2], f2);
    atomic_add_force(F, b[5], f5);
    sdot += value_dot_tangent(f2, q2) + value_dot_tangent(f5, q5);
    const R c1 = (R)2 * ((R)4 * w) * (t1 - ps1), c2 = (R)2 * ((R)4 * w) * (t2 - ps2);
    const int i1[4] = {b[0], b[1], b[3], b[4]}, i2[4] = {b[6], b[4], b[3], b[1]};
    const Pt p1[4] = {q0, q1, q3, q4}, p2[4] = {q6, q4, q3, q1};
    for (int c = 0; c < 4; ++c) {
        const V3<R> fa = g1[c] * c1, fb = g2[c] * c2;
        atomic_add_force(F, i1[c], fa);
        atomic_add_force(F, i2[c], fb);
        sdot += value_dot_tangent(fa, p1[c]) + value_dot_tangent(fb, p2[c]);
    }
    if (vdotf) *vdotf = sdot;
    return e;
}

// Bond or angle g of the concatenated bonded list (polymer_constraints,
// fast.cpp:2354-2520: bead_bond_force k (r - r0)^2, bead_valence_force
// k (theta - theta0)^2): adds the forces, returns the energy in *eb or *ea and the
// end beads (i, k) whose WCA the reference removes; false past the list's end.
template <class R, class L = LdPlain, class FT>
TIS_DEV bool bonded_term(const Dev &D, FT F, int g, int nbond, int nang, int *i_out, int *k_out, double *eb,
                         double *ea) {
    if (g < nbond) {
        const int2 ij = D.bond_ij[g];
        const double2 kr = D.bond_kr[g];
        V3<R> d = diff<R>(PtLoad<R, L>::at(D, ij.x), PtLoad<R, L>::at(D, ij.y));
        R rr;
        R t = bond_coef<R>(d, (R)kr.x, (R)kr.y, &rr);
        atomic_add_force(F, ij.x, d * (-t));
        atomic_add_force(F, ij.y, d * t);
        *eb = kr.x * ((double)rr - kr.y) * ((double)rr - kr.y);
        *i_out = ij.x;
        *k_out = ij.y;
        return true;
    }
    if (g < nbond + nang) {
        const int4 t3 = D.ang_ijk[g - nbond];
        const double2 kt = D.ang_kt[g - nbond];
        const int i = t3.x, j = t3.y, k = t3.z;
        const auto pj = PtLoad<R, L>::at(D, j);
        V3<R> ba = diff<R>(pj, PtLoad<R, L>::at(D, i)), bc = diff<R>(pj, PtLoad<R, L>::at(D, k));
        R c;
        V3<R> g0, g2;
        angle_grad<R>(ba, bc, &c, &g0, &g2);
        c = unit_clamp(c);
        R sn = sqrt((R)1 - c * c), th = acos(c);
        R coef = sn > (R)0 ? (R)-2 * (R)kt.x * ((R)kt.y - th) / sn : (R)0;
        V3<R> f0 = g0 * coef, f2 = g2 * coef;
        atomic_add_force(F, i, f0);
        atomic_add_force(F, k, f2);
        atomic_add_force(F, j, (f0 + f2) * (R)-1);
        *ea = kt.x * ((double)th - kt.y) * ((double)th - kt.y);
        *i_out = i;
        *k_out = k;
        return true;
    }
    return false;
}

// ------------------------------------------------------------ k_force
constexpr int kPairLanes = 4;   // default lanes per particle in the pair gather (D.pair_lanes)
constexpr int kMaxStSlices = 64;  // stack workgroups per replica in a step (D.st_slices)
// ROLE 0 (multi-body): blocks [0, n_rep) stacking of replica b, then the
// backbone terms.  ROLE 1: pair gather.  The two roles are separate kernels so
// that the register-heavy stacking code does not lower the occupancy of the
// latency-bound gather; they run concurrently on two streams.
// eval: kEvalStep (a step: no energies, bonded / 1-3 WCA left out, LR only when
// due), kEvalFull (per-term evaluation with the reference's split: bonded WCA
// added in the pair sum and removed in the bonded terms, energies, LR always),
// kEvalEnergy (as kEvalFull without writing pair forces), kEvalForce (the step's
// forces with a fresh LR term, no energies).
enum { kEvalStep = 0, kEvalFull = 1, kEvalEnergy = 2, kEvalForce = 3 };
template <class R, class L>
TIS_DEV void pair_gather(const Dev &D, int t, int sub, int lanes, bool valid, unsigned mask, int eval);
template <class R, int ROLE>
__global__ void __launch_bounds__(kBlock) k_force(Dev D, int nb_bonded, unsigned mask, int eval) {
    double4 *F = D.f_mb;
    const bool want_e = eval == kEvalFull || eval == kEvalEnergy;
    const bool split = want_e;
    // stack blocks: in a step each replica's stacks are shared out over
    // D.st_slices workgroups (one stack per thread), whose energy sums k_integrate
    // adds up; evaluations use one workgroup per replica
    const int sb = eval == kEvalStep ? D.st_slices : 1;
    if (ROLE == 0 && (int)blockIdx.x < D.n_rep * sb) {
        if (!(mask & 1)) return;
        const int r = blockIdx.x / sb, slice = blockIdx.x % sb;
        const RepParams &P = D.rp[r];
        RepState &S = D.rs[r];
        __shared__ unsigned char off[4096];
        __shared__ double red[8];
        for (int k = threadIdx.x; k < P.nst2; k += blockDim.x) off[k] = 0;
        __syncthreads();
        for (int k = threadIdx.x; k < P.nst3; k += blockDim.x) {
            const int g = P.st30 + k;
            const int *b = D.st3_b + 6 * g;
            V3<R> v = diff<R>(D.pos[b[1]], D.pos[b[4]]);
            R dr = sqrt(dot(v, v)) - (R)D.st3_g[6 * g];
            if (dr < (R)10) {
                const int2 dd = D.st3_dis[g];
                if (dd.x >= 0) off[dd.x] = 1;
                if (dd.y >= 0) off[dd.y] = 1;
            }
        }
        __syncthreads();
        double e2 = 0, e3 = 0;
        // tertiary stacks on whole wavefronts of their own (indices padded to a
        // multiple of 64), so that no wavefront runs both code paths
        const int n3p = (P.nst3 + 63) & ~63;
        for (int k = slice * blockDim.x + threadIdx.x; k < n3p + P.nst2; k += sb * blockDim.x) {
            if (k < P.nst3) {
                const int g = P.st30 + k;
                const double e = (double)tertiary_stack<R>(D, F, g, nullptr);
                D.st_energy[g] = e;
                e3 += e;
            } else if (k >= n3p) {
                const int m = k - n3p;
                if (off[m]) continue;
                e2 += (double)secondary_stack<R>(D, F, P.st20 + m, nullptr);
            }
        }
        double s2 = block_sum_double(e2, red);
        double s3 = block_sum_double(e3, red);
        if (threadIdx.x == 0) {
            if (sb == 1) {
                S.e[EI_E2_ST] = s2;
                S.e[EI_E3_ST] = s3;
            } else {
                D.st_part[2 * (kMaxStSlices * r + slice)] = s2;
                D.st_part[2 * (kMaxStSlices * r + slice) + 1] = s3;
            }
        }
        return;
    }
    int blk = ROLE == 0 ? (int)blockIdx.x - D.n_rep * sb : (int)blockIdx.x;
    if (ROLE == 0) {
        if (blk >= nb_bonded) return;
        if (!(mask & 4)) return;
        const int g = blk * blockDim.x + threadIdx.x;
        const int nbond = D.rp[D.n_rep - 1].bond0 + D.rp[D.n_rep - 1].nbond;
        const int nang = D.rp[D.n_rep - 1].ang0 + D.rp[D.n_rep - 1].nang;
        int i, k;
        double eb = 0, ea = 0;
        if (!bonded_term<R>(D, F, g, nbond, nang, &i, &k, &eb, &ea)) return;
        // remove the WCA force of the bonded / 1-3 pair (reduce_forces, fast.cpp:880-897).
        // The step path leaves that pair's WCA out of the pair sum instead, which is
        // the same force without the cancellation of two terms that diverge as the
        // shifted WCA distance r + 1.5852 - D_LJ goes to 0; a pair listed by several
        // terms is still removed once per extra term.
        if (!split && !D.bsub[g]) return;   // energies are only summed in evaluations
        const int r = D.prep[i];
        const RepParams &P = D.rp[r];
        double3 dd;
        const double r2 = min_image_r2(D.pos[i], D.pos[k], P.L, P.hL, &dd);
        const int ab = (D.ptype[i] & 15) * kNT + (D.ptype[k] & 15);
        const R f = -wca_factor<R>(r2, ab, P);
        V3<R> v = {(R)dd.x * f, (R)dd.y * f, (R)dd.z * f};
        atomic_add_force(F, i, v);
        atomic_add_force(F, k, v * (R)-1);
        if (want_e) {
            RepState &S = D.rs[r];
            unsafeAtomicAdd(&S.e[EI_BOND], eb);
            unsafeAtomicAdd(&S.e[EI_ANGLE], ea);
            unsafeAtomicAdd(&S.e[EI_WCA], -wca_energy(r2, ab, P));
        }
        return;
    }
    // ---- pair gather: D.pair_lanes lanes per particle (cell order; a power of
    // two up to 64), each summing a strided share of the neighbours, then a
    // butterfly reduction inside the group
    const int lanes = D.pair_lanes;
    const int t = (blk * blockDim.x + threadIdx.x) / lanes;
    pair_gather<R, LdPlain>(D, t, threadIdx.x % lanes, lanes, t < D.n_tot, mask, eval);
}

constexpr int kGatherBatch = 2;   // list entries per lane whose loads are issued together
// Pair forces of the particle at cell-order index t (valid: t names a particle;
// every lane of the wavefront takes part, as the lanes of a particle's group
// exchange partial sums), lane sub of its group of `lanes`.
template <class R, class L>
TIS_DEV void pair_gather(const Dev &D, int t, int sub, int lanes, bool valid, unsigned mask, int eval) {
    const bool want_e = eval == kEvalFull || eval == kEvalEnergy;
    const bool split = want_e;
    const bool write_f = eval != kEvalEnergy;
    // own and neighbour positions from the cell-ordered copy (kept current by
    // the integrator): neighbours in the same few cells share cache lines
    const double4 xi = L::ld(D.cpos + (valid ? t : 0));
    const int4 mt = L::ld(D.nb_meta + (valid ? t : 0));   // slot, counts and replica, beside the position
    const int p = cp_index(xi.w);
    const int r = mt.w;
    const RepParams &P = D.rp[r];
    RepState &S = D.rs[r];
    const int pti = cp_type(xi.w);
    const int ti = pti & 15;
    const bool chg_i = (pti >> 8) & 1;
    const int *nb = D.pool + ((size_t)mt.x << 4);
    const int nsr = valid ? mt.y : 0, nlr = valid ? mt.z : 0;
    auto group_sum = [lanes](double v) {
        for (int o = 1; o < lanes; o <<= 1) v += __shfl_xor(v, o, 64);
        return v;
    };
    if (mask & 8) {
        R fx = 0, fy = 0, fz = 0;
        double ew = 0, ec = 0;
        // kGatherBatch entries of this lane at a time: their indices, then their
        // positions, are loaded together (two load levels per batch, not per entry)
        for (int q0 = sub; q0 < nsr; q0 += kGatherBatch * lanes) {
            int eb[kGatherBatch];
            double4 xb[kGatherBatch];
#pragma unroll
            for (int k = 0; k < kGatherBatch; ++k) {
                const int q = q0 + k * lanes;
                eb[k] = q < nsr ? L::ld(nb + q) : 0;
            }
#pragma unroll
            for (int k = 0; k < kGatherBatch; ++k) {
                const int q = q0 + k * lanes;
                xb[k] = q < nsr ? L::ld(D.cpos + (eb[k] < 0 ? ~eb[k] : eb[k])) : xi;
            }
#pragma unroll
            for (int k = 0; k < kGatherBatch; ++k) {
                if (q0 + k * lanes >= nsr) break;
                const int e = eb[k];
                const bool bonded = e < 0;
                const double4 xj = xb[k];
                double3 dd;
                const double r2 = min_image_r2(xi, xj, P.L, P.hL, &dd);
                const int ptj = cp_type(xj.w);
                const int ab = ti * kNT + (ptj & 15);
                // a bonded / 1-3 pair carries no net WCA force (the reference adds it
                // here and removes it in the bonded terms); only evaluations, which
                // report the reference's split by term, compute it
                R f = (bonded && !split) ? (R)0 : wca_factor<R>(r2, ab, P);
                const bool both = chg_i && ((ptj >> 8) & 1) && r2 <= kCoulCut2;
                if (both) {
                    const R rr = sqrt((R)r2);
                    f += (R)P.qlb[ab] / (rr * (R)r2);
                }
                fx += (R)dd.x * f; fy += (R)dd.y * f; fz += (R)dd.z * f;
                if (want_e) {
                    ew += wca_energy(r2, ab, P);
                    if (both) ec += P.qlb[ab] / sqrt(r2);
                }
            }
        }
        const double gx = group_sum((double)fx), gy = group_sum((double)fy), gz = group_sum((double)fz);
        if (valid && sub == 0 && write_f) D.f_sr[p] = make_double4(gx, gy, gz, 0.0);
        if (want_e) {
            ew = group_sum(ew);
            ec = group_sum(ec);
            if (valid && sub == 0) {
                unsafeAtomicAdd(&S.e[EI_WCA], 0.5 * ew);
                unsafeAtomicAdd(&S.e[EI_COUL_SR], 0.5 * ec);
            }
        }
    }
    // the LR refresh is decided per replica; a group never spans two particles,
    // but a wavefront may span two replicas, so the test stays per lane
    const bool lr = (mask & 16) && (eval != kEvalStep || L::ld(&S.lr_due));
    if (__ballot(lr)) {
        R fx = 0, fy = 0, fz = 0;
        double ec = 0;
        const int n_lr = lr ? nlr : 0;
        const int *nb_lr = nb + P.cap_sr;
        for (int q0 = sub; q0 < n_lr; q0 += kGatherBatch * lanes) {
            int eb[kGatherBatch];
            double4 xb[kGatherBatch];
#pragma unroll
            for (int k = 0; k < kGatherBatch; ++k) {
                const int q = q0 + k * lanes;
                eb[k] = q < n_lr ? L::ld(nb_lr + q) : 0;
            }
#pragma unroll
            for (int k = 0; k < kGatherBatch; ++k) {
                const int q = q0 + k * lanes;
                xb[k] = q < n_lr ? L::ld(D.cpos + eb[k]) : xi;
            }
#pragma unroll
            for (int k = 0; k < kGatherBatch; ++k) {
                if (q0 + k * lanes >= n_lr) break;
                const double4 xj = xb[k];
                double3 dd;
                const double r2 = min_image_r2(xi, xj, P.L, P.hL, &dd);
                if (r2 > kCoulCut2) continue;
                const int ab = ti * kNT + (cp_type(xj.w) & 15);
                const R rr = sqrt((R)r2);
                const R f = (R)P.qlb[ab] / (rr * (R)r2);
                fx += (R)dd.x * f; fy += (R)dd.y * f; fz += (R)dd.z * f;
                if (want_e) ec += P.qlb[ab] / sqrt(r2);
            }
        }
        const double gx = group_sum((double)fx), gy = group_sum((double)fy), gz = group_sum((double)fz);
        if (lr && valid && sub == 0 && write_f) D.f_lr[p] = make_double4(gx, gy, gz, 0.0);
        if (want_e) {
            ec = group_sum(ec);
            if (lr && valid && sub == 0) unsafeAtomicAdd(&S.e[EI_COUL_LR], 0.5 * ec);
        }
    }
}

// ------------------------------------------------------------ k_multibody
// The multi-body terms of a step (stacking_interactions fast.cpp:1911-2141,
// polymer_constraints :2354-2520) with each replica's bead forces summed in LDS:
// grid (slices, replicas); slice s owns the terms whose lowest bead lies in its
// share of the replica's beads (term lists sorted by that bead, D.mb_terms, with
// per-slice offsets D.mb_off: [(r (S + 1) + s) 4 + type], types tertiary stack,
// secondary stack, bond, angle).  The secondary stacks, bonds and angles of a
// slice span a few beads past its share, the bead window D.mb_win[2 (r S + s)]
// (first bead, end); their forces are summed in LDS over that window, so a
// bead's contributions meet in one or two workgroups and reach f_mb with one
// atomic per component and workgroup instead of one per term.  Tertiary stacks
// join distant beads and add theirs to f_mb directly.  Dynamic LDS: 3 doubles
// per bead of the largest window.  Energies as k_force in a step: tertiary
// stacks into st_energy, each slice's E2 / E3 stack sums into st_part (summed
// by the step's close).
constexpr int kMbBlock = 256;
// Slice sl of S of replica r, one workgroup (any multiple of 64 threads up to
// 1024); LDS: lf (3 doubles per bead of the window), off (4096 bytes), red (16 doubles).
template <class R, class L>
TIS_DEV void multibody_slice(const Dev &D, int r, int sl, int S, double *lf, unsigned char *off, double *red) {
    const RepParams &P = D.rp[r];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int w_lo = D.mb_win[2 * (r * S + sl)], w_n = D.mb_win[2 * (r * S + sl) + 1] - w_lo;
    for (int k = tid; k < 3 * w_n; k += nt) lf[k] = 0.0;
    for (int k = tid; k < P.nst2; k += nt) off[k] = 0;
    __syncthreads();
    // secondary stacks switched off by formed tertiary stacks (every slice)
    for (int k = tid; k < P.nst3; k += nt) {
        const int g = P.st30 + k;
        const int *b = D.st3_b + 6 * g;
        V3<R> v = diff<R>(L::ld(D.pos + b[1]), L::ld(D.pos + b[4]));
        R dr = sqrt(dot(v, v)) - (R)D.st3_g[6 * g];
        if (dr < (R)10) {
            const int2 dd = D.st3_dis[g];
            if (dd.x >= 0) off[dd.x] = 1;
            if (dd.y >= 0) off[dd.y] = 1;
        }
    }
    __syncthreads();
    const LdsForce F{lf, w_lo};
    const int *o0 = D.mb_off + 4 * ((size_t)r * (S + 1) + sl), *o1 = o0 + 4;
    double e2 = 0.0, e3 = 0.0;
    for (int q = o0[0] + tid; q < o1[0]; q += nt) {
        const int g = D.mb_terms[q];
        const double e = (double)tertiary_stack<R, L>(D, D.f_mb, g, nullptr);
        D.st_energy[g] = e;
        e3 += e;
    }
    for (int q = o0[1] + tid; q < o1[1]; q += nt) {
        const int g = D.mb_terms[q];
        if (!off[g - P.st20]) e2 += (double)secondary_stack<R, L>(D, F, g, nullptr);
    }
    const int nbond = D.rp[D.n_rep - 1].bond0 + D.rp[D.n_rep - 1].nbond;
    const int nang = D.rp[D.n_rep - 1].ang0 + D.rp[D.n_rep - 1].nang;
    const int nbs = o1[2] - o0[2], nts = nbs + o1[3] - o0[3];
    for (int t = tid; t < nts; t += nt) {   // the slice's bonds, then its angles
        const int g = D.mb_terms[t < nbs ? o0[2] + t : o0[3] + t - nbs];
        int i, k;
        double eb, ea;
        bonded_term<R, L>(D, F, g, nbond, nang, &i, &k, &eb, &ea);
        if (!D.bsub[g]) continue;
        // a pair listed by a second bonded term: its WCA removed once more (the
        // step leaves bonded / 1-3 pairs out of the pair sum; k_force explains)
        double3 dd;
        const double r2 = min_image_r2(L::ld(D.pos + i), L::ld(D.pos + k), P.L, P.hL, &dd);
        const int ab = (D.ptype[i] & 15) * kNT + (D.ptype[k] & 15);
        const R f = -wca_factor<R>(r2, ab, P);
        V3<R> v = {(R)dd.x * f, (R)dd.y * f, (R)dd.z * f};
        atomic_add_force(F, i, v);
        atomic_add_force(F, k, v * (R)-1);
    }
    __syncthreads();
    for (int k = tid; k < w_n; k += nt) {
        const double x = lf[3 * k], y = lf[3 * k + 1], z = lf[3 * k + 2];
        if (x != 0.0 || y != 0.0 || z != 0.0) {
            double4 *f = D.f_mb + w_lo + k;
            unsafeAtomicAdd(&f->x, x);
            unsafeAtomicAdd(&f->y, y);
            unsafeAtomicAdd(&f->z, z);
        }
    }
    const double s2 = block_sum_double(e2, red);
    const double s3 = block_sum_double(e3, red);
    if (tid == 0) {
        D.st_part[2 * (kMaxStSlices * r + sl)] = s2;
        D.st_part[2 * (kMaxStSlices * r + sl) + 1] = s3;
    }
}
template <class R>
__global__ void __launch_bounds__(kMbBlock) k_multibody(Dev D) {
    extern __shared__ double lf[];
    __shared__ unsigned char off[4096];
    __shared__ double red[kMbBlock / 64];
    multibody_slice<R, LdPlain>(D, blockIdx.y, blockIdx.x, gridDim.x, lf, off, red);
}

// ------------------------------------------------------------ k_integrate
// klok, epoch: the replica's step counter and Philox epoch.
TIS_DEV void noise3(const Dev &D, int r, const RepParams &P, long long klok, unsigned epoch, int p, double out[3]) {
    const int local = p - P.p0;
    if (D.rng == 0) {
        Philox4 o = philox4x32_10((uint32_t)klok, (uint32_t)(klok >> 32), (uint32_t)local, kTagNoise,
                                  philox_key0(D, r), (uint32_t)(D.seed >> 32) + epoch);
        const double a = sqrt(-2.0 * log(u01_open_closed(o.v[0]))), b = 2.0 * kPi * u01_closed_open(o.v[1]);
        const double c = sqrt(-2.0 * log(u01_open_closed(o.v[2]))), d = 2.0 * kPi * u01_closed_open(o.v[3]);
        out[0] = a * cos(b);
        out[1] = a * sin(b);
        out[2] = c * cos(d);
    } else {
        // particle i (1-based) takes slots i % 1000 of the x, y, z blocks (fast.cpp:2616-2623)
        const int k = (local + 1) % 1000;
        const double *mw = D.maxwell + 3000 * (size_t)r;
        out[0] = mw[k];
        out[1] = mw[k + 1000];
        out[2] = mw[k + 2000];
    }
}

// Random force sigma_type xi of particle p in the current step (full_forces,
// fast.cpp:2598-2624).  With Philox draws it depends only on the step counter,
// the particle and the key, so k_noise computes it beside the step's force
// kernels into D.nz; the reference stream's table is made by the HB resolution
// of the same step, so that mode forms it in the integrator.
template <class L>
TIS_DEV double3 random_force(const Dev &D, int r, const RepParams &P, const RepState &S, int p, int t) {
    if (D.rng == 0) {
        const double4 z = L::ld(D.nz + p);
        return make_double3(z.x, z.y, z.z);
    }
    double nz[3];
    noise3(D, r, P, L::ld(&S.klok), L::ld(&S.rng_epoch), p, nz);
    const double sg = P.sigma[t] * D.noise_scale;
    return make_double3(nz[0] * sg, nz[1] * sg, nz[2] * sg);
}

// k_noise (Philox draws only): D.nz[p] = sigma_type * noise_scale * xi for the
// step about to run, one thread per particle.
// klok, epoch: replica r's step counter and Philox epoch.
TIS_DEV void noise_one(const Dev &D, int r, int p, long long klok, unsigned epoch) {
    const int t = D.ptype[p] & 15;
    const RepParams &P = D.rp[r];
    double nz[3];
    noise3(D, r, P, klok, epoch, p, nz);
    const double sg = P.sigma[t] * D.noise_scale;
    D.nz[p] = make_double4(nz[0] * sg, nz[1] * sg, nz[2] * sg, 0.0);
}
__global__ void __launch_bounds__(kBlock) k_noise(Dev D) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.n_tot) return;
    const int r = D.prep[p];
    noise_one(D, r, p, D.rs[r].klok, D.rs[r].rng_epoch);
}

// leap_frog_move_atom (fast.cpp:570-642) for particle p; returns the new position
// and velocity (position before recentring).  Only RNA beads carry multi-body
// forces (ions' f_mb entries stay zero and are not read).
template <class L>
TIS_DEV void leapfrog(const Dev &D, int r, const RepParams &P, const RepState &S, int p, double xn[3],
                      double vn[3], double xs[3]) {
    const int t = D.ptype[p] & 15;
    const bool bead = p - P.p0 < P.nbeads;
    const double4 f0 = L::ld(D.f_sr + p), f1 = L::ld(D.f_lr + p);
    const double4 f2 = bead ? L::ld(D.f_mb + p) : make_double4(0.0, 0.0, 0.0, 0.0);
    const double3 nz = random_force<L>(D, r, P, S, p, t);
    const double f[3] = {f0.x + f1.x + f2.x + nz.x, f0.y + f1.y + f2.y + nz.y, f0.z + f1.z + f2.z + nz.z};
    const double4 x = L::ld(D.pos + p), xo = L::ld(D.old + p), v = L::ld(D.vel + p);
    const double xv[3] = {x.x, x.y, x.z}, xov[3] = {xo.x, xo.y, xo.z}, vv[3] = {v.x, v.y, v.z};
    for (int c = 0; c < 3; ++c) vn[c] = P.K1[t] * vv[c] + P.K2[t] * f[c];
    double vsq = P.mass[t] * (vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]) / P.vcap;
    if (vsq > 1.0) {
        vsq = 1.0 / sqrt(vsq);
        for (int c = 0; c < 3; ++c) vn[c] *= vsq;
    }
    for (int c = 0; c < 3; ++c) {
        xs[c] = xv[c];
        xn[c] = 2.0 * xv[c] - xov[c] + (vn[c] - vv[c]) * P.h;
    }
}

// The new state goes to the other buffer set (pos_nx, old_nx, vel_nx; every
// read of this step's state is from the current set, so no particle's write can
// reach another's read, the recentring bead's included); the cell-ordered copy
// the pair gather reads is refreshed (whole entries, identity and type word too) and the
// multi-body force buffer of the next step cleared.
// Particle p (valid: p names a particle; every lane of the wavefront takes part
// in the displacement maximum, which goes to word `slot` of its replica).
template <class L>
TIS_DEV void integrate_one(const Dev &D, int p, bool valid, int slot) {
    double d2 = 0.0;   // squared displacement of this thread's particle
    int rp = -1;       // its replica
    if (valid) {
        const int r = D.prep[p];
        const RepParams &P = D.rp[r];
        RepState &S = D.rs[r];
        double xn[3], vn[3], xs[3], kx[3], kv[3], ks[3];
        leapfrog<L>(D, r, P, S, p, xn, vn, xs);
        leapfrog<L>(D, r, P, S, P.kh, kx, kv, ks);   // the recentring bead (adjust_box, fast.cpp:2632-2656)
        double on[3];
        for (int c = 0; c < 3; ++c) {
            const double s = P.hL[c] - kx[c];
            xn[c] += s;
            on[c] = xs[c] + s;
        }
        if (p - P.p0 >= P.nbeads) {
            for (int c = 0; c < 3; ++c) {
                if (xn[c] < 0) { xn[c] += P.L[c]; on[c] += P.L[c]; }
                if (xn[c] > P.L[c]) { xn[c] -= P.L[c]; on[c] -= P.L[c]; }
            }
        }
        const double dx = xn[0] - on[0], dy = xn[1] - on[1], dz = xn[2] - on[2];
        d2 = dx * dx + dy * dy + dz * dz;
        rp = r;
        D.pos_nx[p] = make_double4(xn[0], xn[1], xn[2], 0.0);
        D.old_nx[p] = make_double4(on[0], on[1], on[2], 0.0);
        D.vel_nx[p] = make_double4(vn[0], vn[1], vn[2], 0.0);
        if (p - P.p0 < P.nbeads) D.f_mb_nx[p] = make_double4(0.0, 0.0, 0.0, 0.0);
        // the whole 32-byte entry (identity and type word again), not a partial one
        D.cpos[L::ld(D.crank + p)] = make_double4(xn[0], xn[1], xn[2], cp_pack(p, D.ptype[p]));
    }
    // largest displacement of the step per replica (d2 >= 0, so its bits order
    // like the values): a wavefront maximum when the wavefront lies in one
    // replica, then one atomic per wavefront into one of kMaxSlots words of the
    // replica, each on a cache line of its own (by wavefront index, so no one
    // word or line takes every wavefront's atomic)
    const int r0 = __builtin_amdgcn_readfirstlane(rp);
    if (__ballot(rp == r0) == ~0ull) {
        double m = d2;
        for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
        if ((threadIdx.x & 63) == 0 && r0 >= 0)
            atomicMax(&D.maxd2[((size_t)r0 * kMaxSlots + slot) * kSlotStride], (unsigned long long)__double_as_longlong(m));
    } else if (rp >= 0) {
        atomicMax(&D.maxd2[((size_t)rp * kMaxSlots + slot) * kSlotStride], (unsigned long long)__double_as_longlong(d2));
    }
}

// Closes the step of every replica (a wavefront per replica): the list-rebuild
// decision from the largest displacement k_integrate recorded (check_shifts,
// fast.cpp:1741-1776), the step's stack and HB energy sums, the HB table switch,
// the step counter and the LR refresh flag.  The per-slice, per-slot and
// per-group parts are read by the lanes at once and reduced across the
// wavefront (their sums in the order of the parts, as one thread would add them).
// Replica r, one whole wavefront.
template <class L>
TIS_DEV void commit_replica(const Dev &D, int r) {
    const int lane = threadIdx.x & 63;
    RepState &S = D.rs[r];
    const RepParams &P = D.rp[r];
    const int ns = D.st_slices > 1 ? D.st_slices : 0;
    double s2 = 0.0, s3 = 0.0, h2 = 0.0, h3 = 0.0;
    unsigned long long mb = 0;
    int it = 0;
    if (lane < ns) {
        s2 = L::ld(D.st_part + 2 * (kMaxStSlices * r + lane));
        s3 = L::ld(D.st_part + 2 * (kMaxStSlices * r + lane) + 1);
    }
    if (lane < kMaxSlots) {
        unsigned long long *w = D.maxd2 + ((size_t)r * kMaxSlots + lane) * kSlotStride;
        mb = L::ld(w);
        *w = 0;
    }
    if (lane < D.hb_groups) {
        h2 = L::ld(D.hb_e_part + 2 * (kMaxHbGroups * r + lane));
        h3 = L::ld(D.hb_e_part + 2 * (kMaxHbGroups * r + lane) + 1);
        it = L::ld(D.hb_iter_part + kMaxHbGroups * r + lane);
    }
    // lane 0 adds the parts in order (a readlane each; the sums match a serial loop)
    double t2 = 0.0, t3 = 0.0, u2 = 0.0, u3 = 0.0;
    for (int q = 0; q < ns; ++q) { t2 += readlane_d(s2, q); t3 += readlane_d(s3, q); }
    for (int q = 0; q < D.hb_groups; ++q) { u2 += readlane_d(h2, q); u3 += readlane_d(h3, q); }
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long x = (unsigned long long)__shfl_xor((long long)mb, o, 64);
        mb = mb > x ? mb : x;
        it = max(it, __shfl_xor(it, o, 64));
    }
    if (lane != 0) return;
    if (ns) {
        S.e[EI_E2_ST] = t2;
        S.e[EI_E3_ST] = t3;
    }
    S.e[EI_E2_HB] = u2;
    S.e[EI_E3_HB] = u3;
    S.hb_iters = it;
    const double m = sqrt(__longlong_as_double((long long)mb));
    // a NaN or infinite coordinate (its bits order above every finite
    // displacement) stops the run with an error instead of propagating
    if (!(m < 1e30)) atomicOr(&S.overflow, 32);
    double prog = L::ld(&S.progress) + (m + m);
    const int need = prog > P.skin;
    if (need) { prog = 0; S.n_rebuilds = L::ld(&S.n_rebuilds) + 1; }
    S.progress = prog;
    S.need_rebuild = need;
    S.tab_fresh = 0;
    if (L::ld(&S.tab_pending)) {   // HB tables built beside this step's resolution
        S.tab = L::ld(&S.tab) ^ 1;
        S.n_act = L::ld(&S.nx_act);
        S.n_asite = L::ld(&S.nx_asite);
        S.n_ainc = L::ld(&S.nx_ainc);
        S.tab_pending = 0;
        S.tab_fresh = 1;
    }
    const long long klok = L::ld(&S.klok) + 1;
    S.klok = klok;
    S.lr_due = (klok % P.lr_every) == 1 || prog == 0;
}

// k_integrate: the integration of every particle, then the close of each
// replica's step (commit_replica, check_shifts) by the workgroup that finishes
// the replica's particles last.  Each workgroup, once its waves' displacement
// maxima are in place (the atomics completed: s_waitcnt), counts itself at each
// replica it touched (a returning atomic on D.close_cnt); the one that brings a
// replica's count to its number of workgroups reads that replica's maxima with
// device-scope loads (coherent with the atomics on the same words) and closes
// the step, so no kernel of its own follows.  Every other input of the close
// comes from kernels the launch waited for.
__global__ void __launch_bounds__(kBlock) k_integrate(Dev D) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    integrate_one<LdPlain>(D, p, p < D.n_tot, (int)(p >> 6) % kMaxSlots);
    __shared__ int last[kBlock], n_last;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int pa = blockIdx.x * blockDim.x, pb = min(pa + (int)blockDim.x, D.n_tot) - 1;
        int n = 0;
        for (int r = D.prep[pa]; r <= D.prep[pb]; ++r) {
            const RepParams &P = D.rp[r];
            const int b0 = P.p0 / (int)blockDim.x, b1 = (P.p0 + P.n - 1) / (int)blockDim.x;
            if (atomicAdd(D.close_cnt + (size_t)r * kCloseStride, 1) == b1 - b0) last[n++] = r;
        }
        n_last = n;
    }
    __syncthreads();
    if (threadIdx.x >= 64) return;
    for (int q = 0; q < n_last; ++q) {
        const int r = last[q];
        if (threadIdx.x == 0) D.close_cnt[(size_t)r * kCloseStride] = 0;
        commit_replica<LdCoh>(D, r);
    }
}


// ------------------------------------------------------------ replica exchange
// Helpers of the on-device temperature exchange between neighbouring replicas
// that describe the same system (no counterpart in the reference, which runs
// its replicas as separate jobs).

__global__ void __launch_bounds__(kBlock) k_zero_energies(Dev D) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < D.n_rep)
        for (int q = 0; q < 12; ++q) D.rs[r].e[q] = 0.0;
}

__global__ void __launch_bounds__(kBlock) k_zero_fmb(Dev D) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < D.n_tot) D.f_mb[p] = make_double4(0.0, 0.0, 0.0, 0.0);
}

// Energy sums of the evaluation just done: which 0 = the replica's own
// configuration (non-HB sum and HB sum), 1 = the partner's configuration
// (non-HB sum).  Every term enters with the sign for which F = -grad U.
__global__ void __launch_bounds__(kBlock) k_remd_collect(Dev D, int which) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= D.n_rep) return;
    const double *e = D.rs[r].e;
    const double nonhb = e[EI_E2_ST] + e[EI_E3_ST] + e[EI_WCA] + e[EI_COUL_SR] + e[EI_COUL_LR] + e[EI_BOND] +
                         e[EI_ANGLE];
    if (which == 0) {
        D.remd_e[4 * r] = nonhb;
        D.remd_e[4 * r + 1] = e[EI_E2_HB] + e[EI_E3_HB];
    } else {
        D.remd_e[4 * r + 2] = nonhb;
    }
}

// Marks the exchangeable pairs (r, r + 1) with r % 2 == parity for swapping
// (state 1); the other pairs are idle (0).
__global__ void __launch_bounds__(kBlock) k_remd_mark(Dev D, int parity) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r + 1 < D.n_rep) D.remd_state[r] = (D.remd_ok[r] && (r & 1) == parity) ? 1 : 0;
}

// mode 0: swap the configurations of the pairs in state 1, keeping the old ones
// in the integrator's other buffer set; the configuration moving from slot a to
// slot b has its velocity and x - x_old scaled by sqrt(T_b / T_a).  mode 1:
// restore the pairs in state 2.  Grid (particle blocks, pairs q, pair (2q +
// parity, 2q + parity + 1)); the slots are marked for a list rebuild and their
// walker ids exchanged.  States are only read here.
__global__ void __launch_bounds__(kBlock) k_remd_swap(Dev D, int parity, int mode) {
    const int q = blockIdx.y;
    const int ra = 2 * q + parity, rb = ra + 1;
    if (rb >= D.n_rep) return;
    if (D.remd_state[ra] != (mode == 0 ? 1 : 2)) return;
    const RepParams &A = D.rp[ra], &B = D.rp[rb];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < A.n) {
        const int pa = A.p0 + i, pb = B.p0 + i;
        if (mode == 0) {
            const double4 xa = D.pos[pa], oa = D.old[pa], va = D.vel[pa];
            const double4 xb = D.pos[pb], ob = D.old[pb], vb = D.vel[pb];
            D.pos_nx[pa] = xa; D.old_nx[pa] = oa; D.vel_nx[pa] = va;
            D.pos_nx[pb] = xb; D.old_nx[pb] = ob; D.vel_nx[pb] = vb;
            const double s_ab = sqrt(B.T / A.T), s_ba = sqrt(A.T / B.T);   // a -> b, b -> a
            D.pos[pa] = xb;
            D.old[pa] = make_double4(xb.x - s_ba * (xb.x - ob.x), xb.y - s_ba * (xb.y - ob.y),
                                     xb.z - s_ba * (xb.z - ob.z), 0.0);
            D.vel[pa] = make_double4(s_ba * vb.x, s_ba * vb.y, s_ba * vb.z, 0.0);
            D.pos[pb] = xa;
            D.old[pb] = make_double4(xa.x - s_ab * (xa.x - oa.x), xa.y - s_ab * (xa.y - oa.y),
                                     xa.z - s_ab * (xa.z - oa.z), 0.0);
            D.vel[pb] = make_double4(s_ab * va.x, s_ab * va.y, s_ab * va.z, 0.0);
        } else {
            D.pos[pa] = D.pos_nx[pa]; D.old[pa] = D.old_nx[pa]; D.vel[pa] = D.vel_nx[pa];
            D.pos[pb] = D.pos_nx[pb]; D.old[pb] = D.old_nx[pb]; D.vel[pb] = D.vel_nx[pb];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int w = D.walker[ra];
        D.walker[ra] = D.walker[rb];
        D.walker[rb] = w;
        for (int r : {ra, rb}) {
            RepState &S = D.rs[r];
            S.need_rebuild = 1;
            S.lr_due = 1;
            S.progress = 0;
        }
    }
}

// Metropolis decision for each swapped pair: D as documented in tis_gpu.h,
// accepted when D <= 0 or u < exp(-D) with u a Philox draw keyed by (seed,
// attempt, pair).  Rejected pairs are marked for restoring.
__global__ void __launch_bounds__(kBlock) k_remd_decide(Dev D, long long attempt) {
    const int ra = blockIdx.x * blockDim.x + threadIdx.x, rb = ra + 1;
    if (rb >= D.n_rep || D.remd_state[ra] != 1) return;
    const double ba = 1.0 / D.rp[ra].T, bb = 1.0 / D.rp[rb].T;
    const double *ea = D.remd_e + 4 * ra, *eb = D.remd_e + 4 * rb;
    const double d = ba * (ea[2] - ea[0]) + bb * (eb[2] - eb[0]) + (eb[1] - ea[1]) * (ba - bb);
    const Philox4 u4 = philox4x32_10((uint32_t)attempt, (uint32_t)(attempt >> 32), (uint32_t)ra, kTagSwap,
                                     (uint32_t)D.seed, (uint32_t)(D.seed >> 32));
    const bool acc = d <= 0.0 || u01_closed_open(u4.v[0]) < exp(-d);
    D.remd_d[ra] = d;
    D.remd_stats[2 * ra] += 1;
    if (acc) {
        D.remd_stats[2 * ra + 1] += 1;
        D.remd_state[ra] = 0;
    } else {
        D.remd_state[ra] = 2;
    }
}

__global__ void __launch_bounds__(kBlock) k_clear_rebuild(Dev D) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < D.n_rep) D.rs[r].need_rebuild = 0;
}

}  // namespace tis
