// Force-matching backward pass (no counterpart in the reference).
//
// Loss L = (1/M) sum_i |F_i(x; theta) - F_i^target|^2 over every force
// component of every replica (M components), with F the step's forces
// (bonded / 1-3 pairs without WCA, fresh long-range term, HB after the valence
// resolution, whose outcome is held fixed).  With v = F - F^target:
//   dL/dx     = (2/M) J^T v = (2/M) J v      (J = dF/dx = -Hessian, symmetric)
//   dL/dtheta = (2/M) v . dF/dtheta
// J v is the tangent of the forces along the position tangent v, carried
// through the same templated force routines with the Dual type.  Every term
// whose parameter is a prefactor (HB strength hb_E, stack energy st_E) has forces
// linear in it, so v . dF/dtheta = (v . F_term) / theta.  The ion radii enter the
// WCA contact distance D_LJ = R_a + R_b of every pair involving an ion.
#pragma once
#include "kernels.h"

namespace tis {

// v = F - F^target per particle into D.jv, and sum |v|^2 into *loss.  The forces
// go to forces (double4 per particle) and / or forces3 (packed xyz); either may
// be null.
__global__ void __launch_bounds__(kBlock) k_fm_residual(Dev D, const double *target, double4 *forces,
                                                          double *forces3, double *loss) {
    __shared__ double red[8];
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (p < D.n_tot) {
        const double4 a = D.f_sr[p], b = D.f_lr[p], c = D.f_mb[p];
        const double tx = target[3 * (size_t)p], ty = target[3 * (size_t)p + 1], tz = target[3 * (size_t)p + 2];
        const double fx = a.x + b.x + c.x, fy = a.y + b.y + c.y, fz = a.z + b.z + c.z;
        if (forces) forces[p] = make_double4(fx, fy, fz, 0.0);
        if (forces3) {
            forces3[3 * (size_t)p] = fx;
            forces3[3 * (size_t)p + 1] = fy;
            forces3[3 * (size_t)p + 2] = fz;
        }
        const double vx = fx - tx, vy = fy - ty, vz = fz - tz;
        D.jv[p] = make_double4(vx, vy, vz, 0.0);
        s = vx * vx + vy * vy + vz * vz;
    }
    s = block_sum_double(s, red);
    if (threadIdx.x == 0) unsafeAtomicAdd(loss, s);
}

// Pair terms with derivatives, one thread per particle over its full lists:
// D.jf[p] = tangent of the pair force on p (the multi-body passes add to it
// afterwards).  Also sums, for the radii of Mg, Cl and K (types 6, 7, 8),
// sum_pairs (v_i - v_j) . d_ij df/dD_LJ  (counted once per ordered pair from
// each side: the i side adds v_i . d_ij df/dD).
// kJvpLanes lanes per particle, each a strided share of its neighbours, then a
// butterfly reduction of the tangent force inside the group.
constexpr int kJvpLanes = 8;
__global__ void __launch_bounds__(kBlock) k_jvp_pairs(Dev D, double *grad_ion_r) {
    __shared__ double red[8];
    const int p = (blockIdx.x * blockDim.x + threadIdx.x) / kJvpLanes;
    const int sub = threadIdx.x % kJvpLanes;
    double gr[3] = {0.0, 0.0, 0.0};
    double tf[3] = {0.0, 0.0, 0.0};
    if (p < D.n_tot) {
        const int r = D.prep[p];
        const RepParams &P = D.rp[r];
        const double4 xi = D.pos[p], vi = D.jv[p];
        const int pti = D.ptype[p], ti = pti & 15;
        const bool chg_i = (pti >> 8) & 1;
        const int *nb = D.pool + P.pool0 + D.nb_start[p];
        const int nsr = D.nb_nsr[p], nlr = D.nb_nlr[p];
        Dual fx = 0.0, fy = 0.0, fz = 0.0;
        for (int q = sub; q < nsr + nlr; q += kJvpLanes) {
            const bool is_sr = q < nsr;
            const int e = is_sr ? nb[q] : nb[P.cap_sr + q - nsr];
            const bool bonded = e < 0;
            const int j = D.sorted[bonded ? ~e : e];
            double3 dd;
            const double r2v = min_image_r2(xi, D.pos[j], P.L, P.hL, &dd);
            const double4 vj = D.jv[j];
            const double tx = vi.x - vj.x, ty = vi.y - vj.y, tz = vi.z - vj.z;
            const Dual dx(dd.x, tx), dy(dd.y, ty), dz(dd.z, tz);
            const Dual r2 = dx * dx + dy * dy + dz * dz;
            const int ptj = D.ptype[j], tj = ptj & 15;
            const int ab = ti * kNT + tj;
            Dual f = 0.0;
            if (is_sr && !bonded && r2v <= P.d2lj[ab]) {
                const Dual rr = sqrt(r2);
                const Dual rho = rr + kSigLJ - P.dlj[ab];
                const Dual u = Dual(kSigLJ) / rho;
                const Dual u2 = u * u, u4 = u2 * u2, u8 = u4 * u4, u14 = u8 * u4 * u2;
                f = f + Dual(P.flj[ab]) * (u14 - u8) * rho / rr;
                const bool ion_pair = !(ti <= 5 && tj <= 5);
                if (ion_pair) {
                    // df/dD_LJ = F_LJ (13 u^14 - 7 u^8) / r at fixed r
                    const double dfd = P.flj[ab] * (13.0 * u14.v - 7.0 * u8.v) / rr.v;
                    const double c = dfd * (vi.x * dd.x + vi.y * dd.y + vi.z * dd.z);
                    if (ti >= 6) gr[ti - 6] += c;
                    if (tj >= 6) gr[tj - 6] += c;
                }
            }
            if (chg_i && ((ptj >> 8) & 1) && r2v <= kCoulCut2) {
                const Dual rr = sqrt(r2);
                f = f + Dual(P.qlb[ab]) / (rr * r2);
            }
            fx += f * dx;
            fy += f * dy;
            fz += f * dz;
        }
        tf[0] = fx.d;
        tf[1] = fy.d;
        tf[2] = fz.d;
    }
#pragma unroll
    for (int t = 0; t < 3; ++t)
        for (int o = 1; o < kJvpLanes; o <<= 1) tf[t] += __shfl_xor(tf[t], o, 64);
    if (p < D.n_tot && sub == 0) D.jf[p] = make_double4(tf[0], tf[1], tf[2], 0.0);
    for (int t = 0; t < 3; ++t) {
        const double s = block_sum_double(gr[t], red);
        if (threadIdx.x == 0 && s != 0.0) unsafeAtomicAdd(&grad_ion_r[t], s);
    }
}

// Per-term gradient value of term g of replica r: stored at its global index, or
// (local) added at its index within the replica, which sums replicas of one system.
TIS_DEV void put_term_grad(double *dst, int g, int g0, bool local, double v) {
    if (local) unsafeAtomicAdd(&dst[g - g0], v);
    else dst[g] = v;
}

// Stacks (one workgroup per replica, after the tertiary-formation flags of the
// step) and bonded terms (one thread per term) with derivatives: tangents into
// D.jf; per stack, v . F_stack / E_stack into g_st3 / g_st2 (global stack index,
// or summed at the replica-local index when local).
__global__ void __launch_bounds__(kBlock) k_jvp_multibody(Dev D, int nb_bonded, double *g_st3, double *g_st2,
                                                            int local) {
    if ((int)blockIdx.x < D.n_rep) {
        const int r = blockIdx.x;
        const RepParams &P = D.rp[r];
        __shared__ unsigned char off[4096];
        for (int k = threadIdx.x; k < P.nst2; k += blockDim.x) off[k] = 0;
        __syncthreads();
        for (int k = threadIdx.x; k < P.nst3; k += blockDim.x) {
            const int g = P.st30 + k;
            const int *b = D.st3_b + 6 * g;
            V3<double> v = diff<double>(D.pos[b[1]], D.pos[b[4]]);
            if (sqrt(dot(v, v)) - D.st3_g[6 * g] < 10.0) {
                const int2 dd = D.st3_dis[g];
                if (dd.x >= 0) off[dd.x] = 1;
                if (dd.y >= 0) off[dd.y] = 1;
            }
        }
        __syncthreads();
        for (int k = threadIdx.x; k < P.nst3 + P.nst2; k += blockDim.x) {
            double vd = 0.0;
            if (k < P.nst3) {
                const int g = P.st30 + k;
                tertiary_stack<Dual>(D, D.jf, g, &vd);
                put_term_grad(g_st3, g, P.st30, local, vd / D.st3_E[g]);
            } else {
                const int m = k - P.nst3, g = P.st20 + m;
                if (off[m]) { put_term_grad(g_st2, g, P.st20, local, 0.0); continue; }
                secondary_stack<Dual>(D, D.jf, g, &vd);
                put_term_grad(g_st2, g, P.st20, local, vd / D.st2_rE[g].y);
            }
        }
        return;
    }
    const int blk = (int)blockIdx.x - D.n_rep;
    if (blk >= nb_bonded) return;
    const int g = blk * blockDim.x + threadIdx.x;
    const int nbond = D.rp[D.n_rep - 1].bond0 + D.rp[D.n_rep - 1].nbond;
    const int nang = D.rp[D.n_rep - 1].ang0 + D.rp[D.n_rep - 1].nang;
    int i, k;
    double eb, ea;
    if (!bonded_term<Dual>(D, D.jf, g, nbond, nang, &i, &k, &eb, &ea)) return;
    if (!D.bsub[g]) return;
    // a pair listed by a second bonded term: its WCA is removed once more
    const int r = D.prep[i];
    const RepParams &P = D.rp[r];
    double3 dd;
    const double r2v = min_image_r2(D.pos[i], D.pos[k], P.L, P.hL, &dd);
    const int ab = (D.ptype[i] & 15) * kNT + (D.ptype[k] & 15);
    if (r2v > P.d2lj[ab]) return;
    const double4 vi = D.jv[i], vk = D.jv[k];
    const Dual dx(dd.x, vi.x - vk.x), dy(dd.y, vi.y - vk.y), dz(dd.z, vi.z - vk.z);
    const Dual rr = sqrt(dx * dx + dy * dy + dz * dz);
    const Dual rho = rr + kSigLJ - P.dlj[ab];
    const Dual u = Dual(kSigLJ) / rho;
    const Dual u2 = u * u, u4 = u2 * u2, u8 = u4 * u4, u14 = u8 * u4 * u2;
    const Dual f = -(Dual(P.flj[ab]) * (u14 - u8) * rho / rr);
    const V3<Dual> fv = {f * dx, f * dy, f * dz};
    atomic_add_force(D.jf, i, fv);
    atomic_add_force(D.jf, k, fv * Dual(-1.0));
}

// HB candidates kept by the resolution, with derivatives: tangents into D.jf and
// v . F_hb / hb_E into g_hb (global candidate index, or summed at the
// replica-local index when local).  Grid (kActMax / kBlock, replicas).
__global__ void __launch_bounds__(kBlock) k_jvp_hb(Dev D, double *g_hb, int local) {
    const int r = blockIdx.y;
    const RepState &S = D.rs[r];
    const RepParams &P = D.rp[r];
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if ((S.overflow & 2) || a >= S.n_act) return;
    if (D.act_flag[P.hb0 + a] != FLAG_LIVE) return;
    const int g = hb_tab(D, S.tab).act_hb[P.hb0 + a];
    Dual q, psi0[3];
    hb_geometry<Dual>(D, g, &q, psi0, nullptr);
    const Dual e = Dual(D.hb_E[g]) * exp(-q);
    double vd = 0.0;
    sextet_forces<Dual>(D, D.jf, D.hb_b + 6 * g, D.hb_g + 6 * g, -e, psi0, Dual(5.0), Dual(1.5), Dual(0.15), &vd);
    put_term_grad(g_hb, g, P.hb0, local, vd / D.hb_E[g]);
}

}  // namespace tis

namespace tis {

// xyz of n double4 values, times `scale`, into a packed double[3 n] (the layout
// the C ABI hands back).
__global__ void __launch_bounds__(kBlock) k_pack3(const double4 *src, int n, double scale, double *dst) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double4 v = src[p];
    dst[3 * (size_t)p] = scale * v.x;
    dst[3 * (size_t)p + 1] = scale * v.y;
    dst[3 * (size_t)p + 2] = scale * v.z;
}

// dst[k] = scale * sum_r src[r * n + k]: per-term gradients of replicas of one
// system, summed over the replicas on the device.
__global__ void __launch_bounds__(kBlock) k_sum_replicas(const double *src, int n, int n_rep, double scale,
                                                          double *dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double s = 0.0;
    for (int r = 0; r < n_rep; ++r) s += src[(size_t)r * n + k];
    dst[k] = scale * s;
}

__global__ void __launch_bounds__(kBlock) k_fill_zero(double *dst, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = 0.0;
}

// Outputs of the allocation-free force matching from its working sums: s[0] sum
// |v|^2, s[1..3] ion-radius sums, and the per-term sums shb, sst3, sst2 of the n_hb,
// n_st3, n_st2 terms of one replica; the gradients are scaled by sc = 2 / M and
// the loss by 1 / M.  Null outputs are skipped.
__global__ void __launch_bounds__(kBlock) k_fm_finish(const double *s, const double *shb, const double *sst3,
                                                       const double *sst2, int n_hb, int n_st3, int n_st2, double sc,
                                                       double inv_m, double *loss, double *g_ion, double *g_hb,
                                                       double *g_st3, double *g_st2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0 && loss) loss[0] = s[0] * inv_m;
    if (k < 3 && g_ion) g_ion[k] = sc * s[1 + k];
    if (k < n_hb && g_hb) g_hb[k] = sc * shb[k];
    if (k < n_st3 && g_st3) g_st3[k] = sc * sst3[k];
    if (k < n_st2 && g_st2) g_st2[k] = sc * sst2[k];
}

// Parameters of replica 0's region copied to every replica (replicas of one
// system): dst[(r * n + k) * stride + at] = src[k].
__global__ void __launch_bounds__(kBlock) k_set_param(double *dst, int n, int n_rep, int stride, int at,
                                                       const double *src) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * n_rep) return;
    dst[i * stride + at] = src[i % n];
}

}  // namespace tis
