// Equilibrium observables accumulated on the device, one sample per call of
// launch_observe (engine.hip), over every replica of the batch.
//
// The reference writes raw frames (Coordinates, Energies, Contacts files,
// fast.cpp:6276-6316) and computes its observables offline from them:
//   - phosphate-ion contacts: ions of each kind within a cut of each P bead,
//     minimum image in the periodic box (analysis/mg_correlation/mg_bound.F90:97-126),
//     the fraction of frames with at least one Mg bound per P
//     (analysis/mg_correlation/count_mg_bound.py:8-15), and the number of distinct
//     ions within the cut of any P (count_bound_total in
//     analysis/salt_nt/dcd_traj_SaltCount_lB.py:52-62, cut r_kT = 7.2684 A);
//   - chain shape: gyration tensor T = sum_ij (x_i - x_j)(x_i - x_j)^T / (2 N^2)
//     of the beads, Rg = sqrt(tr T), asphericity D and shape parameter S from
//     its eigenvalues (analysis/shape/dcd_traj_shape.F90:193-226);
//   - HB and stack occupancy: per-candidate outcome of the HB valence
//     resolution and per-stack tertiary energies (the Contacts columns).
// Here the same quantities are summed on the device every sample, so a long run
// needs no frames on disk.
//
// Accumulators (zeroed by tis_observe):
//   obs_rep  kObsRep doubles per replica: [0] samples, [1] sum Rg, [2] sum Rg^2,
//            [3] sum D, [4] sum S, [5..13] sums of the nine energy terms in
//            RepState::e order, [14..16] sums of the bound Mg, Cl, K counts
//   obs_p    per particle (P beads only): sums of the Mg, Cl, K counts within the
//            cut and of [Mg count >= 1]
//   obs_hb   per HB candidate: samples in which it was kept by the resolution
//   obs_st3  per tertiary stack: sum of its energy
//   obs_cnt  per replica, 4 ints: bound-ion counts of the current sample (scratch,
//            cleared by k_obs_reduce)
//   obs_keep the step's energies (RepState::e, hb_iters), per-stack energies and HB
//            outcomes, saved before a sample's evaluation and restored after it
#pragma once
#include "kernels.h"

namespace tis {

constexpr int kObsRep = 24;
constexpr int kObsKeep = 13;   // doubles per replica in obs_keep

// Contact counts, one thread per particle.  A P bead counts the ions among its
// SR neighbours within the cut; an ion marks itself bound when a P bead among
// its SR neighbours is within the cut.  The SR lists hold every pair that was
// within r_sr at the last build; since then no pair distance has changed by
// more than the skin plus two steps' displacement, so for cut <= r_sr - skin - 1
// (checked on the host) every pair within the cut is listed.
__global__ void __launch_bounds__(kBlock) k_obs_contacts(Dev D, double cut2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.n_tot) return;
    const int cls = (D.ptype[p] >> 4) & 15;
    if (cls != 0 && (cls < 3 || cls > 5)) return;
    const int r = D.prep[p];
    const RepParams &P = D.rp[r];
    const double4 xi = D.pos[p];
    const int *nb = D.pool + P.pool0 + D.nb_start[p];
    const int nsr = D.nb_nsr[p];
    if (cls == 0) {
        int cnt[3] = {0, 0, 0};
        for (int q = 0; q < nsr; ++q) {
            const int e = nb[q];
            if (e < 0) continue;                       // bonded / 1-3 partner: a bead
            const double4 xj = D.cpos[e];              // entries are cell-order indices
            const int cj = (cp_type(xj.w) >> 4) & 15;
            if (cj < 3 || cj > 5) continue;
            if (min_image_r2(xi, xj, P.L, P.hL) <= cut2) ++cnt[cj - 3];
        }
        double4 a = D.obs_p[p];
        a.x += cnt[0]; a.y += cnt[1]; a.z += cnt[2]; a.w += cnt[0] > 0 ? 1.0 : 0.0;
        D.obs_p[p] = a;
    } else {
        for (int q = 0; q < nsr; ++q) {
            const int e = nb[q];
            if (e < 0) continue;
            const double4 xj = D.cpos[e];
            if (((cp_type(xj.w) >> 4) & 15) != 0) continue;
            if (min_image_r2(xi, xj, P.L, P.hL) <= cut2) {
                atomicAdd(&D.obs_cnt[4 * r + cls - 3], 1);
                break;
            }
        }
    }
}

// Per replica (one workgroup each): shape of the chain, energy terms, bound
// totals, HB and stack occupancy.  The gyration tensor of the pair formula
// equals the covariance (1/N) sum y y^T - ybar ybar^T with y = x - x_ref (x_ref
// the recentring bead, which keeps the moments small).  D and S need only
// invariants of A = T - (tr T / 3) I:  sum_k (w_k - wbar)^2 = tr(A^2) and
// prod_k (w_k - wbar) = det A, so
//   D = 1.5 tr(A^2) / tr(T)^2,   S = 27 det A / tr(T)^3.
__global__ void __launch_bounds__(kBlock) k_obs_reduce(Dev D) {
    const int r = blockIdx.x;
    const RepParams &P = D.rp[r];
    const RepState &S = D.rs[r];
    __shared__ double red[8];
    const double4 ref = D.pos[P.kh];
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = threadIdx.x; k < P.nbeads; k += blockDim.x) {
        const double4 q = D.pos[P.p0 + k];
        const double y0 = q.x - ref.x, y1 = q.y - ref.y, y2 = q.z - ref.z;
        m[0] += y0; m[1] += y1; m[2] += y2;
        m[3] += y0 * y0; m[4] += y1 * y1; m[5] += y2 * y2;
        m[6] += y0 * y1; m[7] += y0 * y2; m[8] += y1 * y2;
    }
    double s[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) s[c] = block_sum_double(m[c], red);
    if (threadIdx.x == 0) {
        const double n = (double)P.nbeads;
        const double c0 = s[0] / n, c1 = s[1] / n, c2 = s[2] / n;
        const double txx = s[3] / n - c0 * c0, tyy = s[4] / n - c1 * c1, tzz = s[5] / n - c2 * c2;
        const double txy = s[6] / n - c0 * c1, txz = s[7] / n - c0 * c2, tyz = s[8] / n - c1 * c2;
        const double tr = txx + tyy + tzz, q3 = tr / 3.0;
        const double a0 = txx - q3, a1 = tyy - q3, a2 = tzz - q3;
        const double trA2 = a0 * a0 + a1 * a1 + a2 * a2 + 2.0 * (txy * txy + txz * txz + tyz * tyz);
        const double detA = a0 * (a1 * a2 - tyz * tyz) - txy * (txy * a2 - tyz * txz) + txz * (txy * tyz - a1 * txz);
        double *o = D.obs_rep + (size_t)kObsRep * r;
        o[0] += 1.0;
        o[1] += sqrt(tr);
        o[2] += tr;
        o[3] += 1.5 * trA2 / (tr * tr);
        o[4] += 27.0 * detA / (tr * tr * tr);
        for (int k = 0; k < 9; ++k) o[5 + k] += S.e[k];
        for (int k = 0; k < 3; ++k) {
            o[14 + k] += (double)D.obs_cnt[4 * r + k];
            D.obs_cnt[4 * r + k] = 0;
        }
    }
    for (int k = threadIdx.x; k < P.nhb; k += blockDim.x) {
        const int g = P.hb0 + k;
        if (D.hb_out[g] == 1) D.obs_hb[g] += 1.0;
    }
    for (int k = threadIdx.x; k < P.nst3; k += blockDim.x) {
        const int g = P.st30 + k;
        D.obs_st3[g] += D.st_energy[g];
    }
}

// Saves (restore = 0) or restores (restore = 1) what the last step's force call
// left for the output files and the outcome report: the energy terms and serial
// depth of each replica, the tertiary stack energies (Contacts columns) and the
// HB outcomes.  A sample evaluates every term afresh at the current positions,
// which would otherwise replace them.  Grid: one workgroup per replica.
__global__ void __launch_bounds__(kBlock) k_obs_keep(Dev D, int restore) {
    const int r = blockIdx.x;
    const RepParams &P = D.rp[r];
    RepState &S = D.rs[r];
    double *ke = D.obs_keep + (size_t)kObsKeep * r;
    if (threadIdx.x < 12) {
        if (restore) S.e[threadIdx.x] = ke[threadIdx.x];
        else ke[threadIdx.x] = S.e[threadIdx.x];
    }
    if (threadIdx.x == 12) {
        if (restore) S.hb_iters = (int)ke[12];
        else ke[12] = (double)S.hb_iters;
    }
    double *kst = D.obs_keep + (size_t)kObsKeep * D.n_rep;
    for (int k = threadIdx.x; k < P.nst3; k += blockDim.x) {
        const int g = P.st30 + k;
        if (restore) D.st_energy[g] = kst[g];
        else kst[g] = D.st_energy[g];
    }
    for (int k = threadIdx.x; k < P.nhb; k += blockDim.x) {
        const int g = P.hb0 + k;
        if (restore) D.hb_out[g] = D.obs_keep_hb[g];
        else D.obs_keep_hb[g] = D.hb_out[g];
    }
}

}  // namespace tis
