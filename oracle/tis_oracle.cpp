// FP64 CPU restatement of the reference's Langevin step (test infrastructure
// only; see tis_oracle.h).  Reference = src/16SCD_Mg_Cl_K_COPY_fast.cpp.
//
// Particle indices are 0-based here; where the reference's arithmetic depends
// on its 1-based numbering (cell index by ceil, noise slot i % 1000) the
// 1-based value is formed explicitly.
#include "tis_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846264338328;
constexpr double kSig = 1.5852;          // WCA shift length (fast.cpp:801-821)
constexpr double kCoulCut2 = 50.0 * 50.0; // D2_CT_CUTOFF (fast.cpp:348)
constexpr double kRandMax = 2147483647.0;

struct V3 { double x, y, z; };
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 scale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }

// Pair parameters for two maxi_key types (read_maxi_key, fast.cpp:2913-2944).
struct PairTab {
    double d_lj[9][9], d2_lj[9][9], d2_min[9][9], e_lj[9][9], f_lj[9][9], q_c[9][9];
};

}  // namespace

struct oracle_ctx {
    // ---- system (copied)
    int N = 0, NB = 0, NU = 0, n3 = 0, n2 = 0, nh = 0, ns = 0;
    double L[3]{}, hL[3]{}, T = 0, lB = 0, h = 0;
    int kh = 0;
    std::vector<int> type, cls, u0, u1;
    std::vector<char> ukind;
    double mass[9]{}, K1[9]{}, K2[9]{}, sigma[9]{};
    PairTab pt{};
    std::vector<int> st3b, st3face, st2b, st_excess;
    std::vector<double> st3g, st3E, st2r, st2E;
    std::vector<int> hbb, hbdode, hbso, hbsi, shbo, shbi, valence;
    std::vector<double> hbg, hbE;
    // ---- lists / settings
    double r_sr2 = 0, r_lr2 = 0, r_cell = 0, skin = 3.0;
    int lr_every = 10;
    std::vector<int> list_sr, list_lr;   // flat (i,j) in the reference's visiting order
    std::vector<int> hb_status, hb_excess0, hb_outcome;
    // ---- state
    std::vector<double> x, xo, v, f, fl;
    long long klok = 1, rebuilds = 0, hb_iters = 0;
    double progress = 0;
    std::vector<int32_t> stream;
    long long cursor = 0;
    bool stream_overflow = false;
    // ---- per-evaluation scratch
    std::vector<double> hb_energy, st_energy;
    std::vector<double> hb_psi0;  // 3 per HB
    std::vector<int> excess;
    double E2_HB = 0, E3_HB = 0, E2_ST = 0, E3_ST = 0;

    V3 p(int i) const { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }
    void addf(std::vector<double> &F, int i, V3 a) {
        F[3 * i] += a.x; F[3 * i + 1] += a.y; F[3 * i + 2] += a.z;
    }

    double uniform() {
        if (cursor >= (long long)stream.size()) { stream_overflow = true; return 0.5; }
        return (double)stream[cursor++] / kRandMax;
    }

    // half_shift (fast.cpp:646-653) applied to x_i - x_j.
    V3 pair_vec(int i, int j) const {
        double d[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
        for (int c = 0; c < 3; ++c) {
            if (d[c] > hL[c]) d[c] = d[c] - L[c];
            if (d[c] < -hL[c]) d[c] = d[c] + L[c];
        }
        return {d[0], d[1], d[2]};
    }

    // LJ_FORCE (fast.cpp:801-821): factor multiplying the pair vector.
    double lj_force(double r2, int a, int b) const {
        if (r2 > pt.d2_lj[a][b]) return 0;
        double r = std::sqrt(r2);
        double rho = r + kSig - pt.d_lj[a][b];
        double u = kSig / rho;
        double u2 = u * u, u4 = u2 * u2, u8 = u4 * u4, u14 = u8 * u4 * u2;
        return pt.f_lj[a][b] * (u14 - u8) * rho / r;
    }
    // WCA_ENERGY (fast.cpp:749-771), with its 1001*E_LJ clamp at very short range.
    double wca_energy(double r2, int a, int b) const {
        if (r2 > pt.d2_lj[a][b]) return 0;
        if (r2 < pt.d2_min[a][b]) return pt.e_lj[a][b] * 1001.0;
        double rho = std::sqrt(r2) + kSig - pt.d_lj[a][b];
        double u = kSig / rho;
        double u2 = u * u, u4 = u2 * u2, u6 = u4 * u2, u12 = u6 * u6;
        return pt.e_lj[a][b] * (u12 - 2.0 * u6 + 1.0);
    }
    bool charged(int i) const { return cls[i] != 1 && cls[i] != 2; }

    // ------------------------------------------------------------ lists
    // separate_in_cells + populate_lists (fast.cpp:1461-1737): cells of at least
    // r_cell, a 13-cell half shell plus the cell itself, in the reference's order.
    void build_lists_cells() {
        int M[3];
        double lc[3];
        for (int c = 0; c < 3; ++c) { M[c] = (int)(L[c] / r_cell); lc[c] = L[c] / M[c]; }
        int MT = M[0] * M[1] * M[2];
        std::vector<std::vector<int>> cell(MT + 1);
        for (int i = 0; i < N; ++i) {
            int ci[3];
            for (int c = 0; c < 3; ++c) {
                double t = x[3 * i + c];
                if (t < 0) t += L[c];
                else if (t > L[c]) t -= L[c];
                ci[c] = (int)std::ceil(t / lc[c]);
                if (ci[c] < 1) ci[c] = 1;   // a coordinate of exactly 0 joins cell 1
                if (ci[c] > M[c]) ci[c] = M[c];
            }
            int m = (ci[0] - 1) * M[1] * M[2] + (ci[1] - 1) * M[2] + ci[2];
            cell[m].push_back(i);
        }
        static const int off[13][3] = {{-1, -1, -1}, {-1, -1, 0}, {-1, -1, 1}, {-1, 0, -1}, {-1, 0, 0},
                                       {-1, 0, 1},   {-1, 1, -1}, {-1, 1, 0},  {-1, 1, 1},  {0, -1, -1},
                                       {0, -1, 0},   {0, -1, 1},  {0, 0, -1}};
        list_sr.clear();
        list_lr.clear();
        auto test_pair = [&](int i1, int i2) {
            V3 d = pair_vec(i1, i2);
            double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
            if (r2 <= r_sr2) {
                list_sr.push_back(i1); list_sr.push_back(i2);
            } else if (r2 <= r_lr2 && charged(i1) && charged(i2)) {
                list_lr.push_back(i1); list_lr.push_back(i2);
            }
        };
        for (int m1 = 1; m1 <= M[0]; ++m1)
            for (int m2 = 1; m2 <= M[1]; ++m2)
                for (int m3 = 1; m3 <= M[2]; ++m3) {
                    int J1 = (m1 - 1) * M[1] * M[2] + (m2 - 1) * M[2] + m3;
                    int mm[3] = {m1, m2, m3};
                    for (int k = 0; k < 13; ++k) {
                        int nn[3];
                        bool skip = false;
                        for (int c = 0; c < 3 && !skip; ++c) {
                            nn[c] = mm[c] + off[k][c];
                            // periodic wrap; with <= 2 cells a wrapped neighbour may
                            // coincide with a cell already visited, which is skipped
                            if (nn[c] > M[c]) { nn[c] -= M[c]; if (nn[c] == mm[c] || nn[c] == mm[c] - 1) skip = true; }
                            else if (nn[c] < 1) { nn[c] += M[c]; if (nn[c] == mm[c] || nn[c] == mm[c] + 1) skip = true; }
                        }
                        if (skip) continue;
                        int J2 = (nn[0] - 1) * M[1] * M[2] + (nn[1] - 1) * M[2] + nn[2];
                        for (int i1 : cell[J1])
                            for (int i2 : cell[J2]) test_pair(i1, i2);
                    }
                    const std::vector<int> &c1 = cell[J1];
                    for (size_t a = 0; a < c1.size(); ++a)
                        for (size_t b = a + 1; b < c1.size(); ++b) test_pair(c1[a], c1[b]);
                }
        after_lists();
    }

    // Same sets by brute force over all pairs (order i<j).
    void build_lists_brute() {
        list_sr.clear();
        list_lr.clear();
        for (int i = 0; i < N; ++i)
            for (int j = i + 1; j < N; ++j) {
                V3 d = pair_vec(i, j);
                double r2 = d.x * d.x + d.y * d.y + d.z * d.z;
                if (r2 <= r_sr2) { list_sr.push_back(i); list_sr.push_back(j); }
                else if (r2 <= r_lr2 && charged(i) && charged(j)) { list_lr.push_back(i); list_lr.push_back(j); }
            }
        after_lists();
    }

    // HB candidate status and HB_EXCESS from the SR list (fast.cpp:1702-1736):
    // a candidate is live (0) iff its (k11, k21) bead pair is an SR pair.
    void after_lists() {
        std::fill(hb_status.begin(), hb_status.end(), -1);
        // index candidates by their bead pair
        std::vector<std::pair<long long, int>> key(nh);
        for (int k = 0; k < nh; ++k) {
            int a = hbb[6 * k + 1], b = hbb[6 * k + 4];
            if (a > b) std::swap(a, b);
            key[k] = {(long long)a * N + b, k};
        }
        std::sort(key.begin(), key.end());
        for (size_t q = 0; q < list_sr.size(); q += 2) {
            int a = list_sr[q], b = list_sr[q + 1];
            if (cls[a] > 2 || cls[b] > 2) continue;
            if (a > b) std::swap(a, b);
            long long kk = (long long)a * N + b;
            auto it = std::lower_bound(key.begin(), key.end(), std::make_pair(kk, -1));
            for (; it != key.end() && it->first == kk; ++it) hb_status[it->second] = 0;
        }
        for (int s = 0; s < ns; ++s) hb_excess0[s] = (shbo[s + 1] - shbo[s]) - valence[s];
        for (int k = 0; k < nh; ++k) {
            if (hb_status[k] == 0) continue;
            for (int q = hbso[k]; q < hbso[k + 1]; ++q) hb_excess0[hbsi[q]] -= 1;
        }
    }

    // ------------------------------------------------------------ geometry
    // valence_angle (fast.cpp:986-1027)
    double angle(int i1, int i2, int i3) const {
        V3 a = sub(p(i1), p(i2)), b = sub(p(i3), p(i2));
        double c = dot(a, b) / std::sqrt(dot(a, a) * dot(b, b));
        return std::acos(c);
    }
    // dihedral_angle (fast.cpp:1255-1302): atan2((b1.n)|b2|, m.n), m = b1 x b2, n = b2 x b3.
    double dihedral(int i1, int i2, int i3, int i4) const {
        V3 b1 = sub(p(i2), p(i1)), b2 = sub(p(i3), p(i2)), b3 = sub(p(i4), p(i3));
        V3 m = cross(b1, b2), n = cross(b2, b3);
        return std::atan2(dot(b1, n) * std::sqrt(dot(b2, b2)), dot(m, n));
    }
    double dist(int i1, int i2) const { V3 d = sub(p(i1), p(i2)); return std::sqrt(dot(d, d)); }

    // Harmonic bond k (r - r0)^2, no periodic image (stacking_bond_force, fast.cpp:957-982).
    void bond_force(std::vector<double> &F, int i1, int i2, double k, double r0) {
        V3 d = sub(p(i2), p(i1));
        double r = std::sqrt(dot(d, d));
        double t = 2.0 * k * (r0 - r) / r;
        addf(F, i1, scale(d, -t));
        addf(F, i2, scale(d, t));
    }
    // Harmonic angle k (theta - theta0)^2 (bead_/stacking_valence_force,
    // fast.cpp:1031-1251); guard = the stacking variant's zero force near pi.
    void angle_force(std::vector<double> &F, int i1, int i2, int i3, double k, double t0, bool guard) {
        V3 a = sub(p(i1), p(i2)), b = sub(p(i3), p(i2));
        double ia = 1.0 / dot(a, a), ib = 1.0 / dot(b, b);
        double iab = std::sqrt(ia * ib);
        double c = dot(a, b) * iab;
        double s, th;
        if (guard && c < -0.99995) { s = 0; th = kPi; }
        else { s = std::sqrt(1.0 - c * c); th = std::acos(c); }
        if (s == 0) return;
        double coef = -2.0 * k * (t0 - th) / s;   // dE/d(cos theta)
        // d cos / d x1 = b/(|a||b|) - c a/|a|^2 ; d cos / d x3 = a/(|a||b|) - c b/|b|^2
        V3 g1 = sub(scale(b, iab), scale(a, c * ia));
        V3 g3 = sub(scale(a, iab), scale(b, c * ib));
        V3 f1 = scale(g1, coef), f3 = scale(g3, coef);
        addf(F, i1, f1);
        addf(F, i3, f3);
        addf(F, i2, scale({f1.x + f3.x, f1.y + f3.y, f1.z + f3.z}, -1.0));
    }
    // Harmonic dihedral k (psi - psi0)^2 (stacking_dihedral_force, fast.cpp:1306-1454).
    // Gradient of psi by the singularity-free cross-product form:
    //   dpsi/dx1 = -|b2| m/|m|^2, dpsi/dx4 = |b2| n/|n|^2,
    //   dpsi/dx2 = -(1 + b1.b2/|b2|^2) dpsi/dx1 + (b3.b2/|b2|^2) dpsi/dx4,
    //   dpsi/dx3 = -(1 + b3.b2/|b2|^2) dpsi/dx4 + (b1.b2/|b2|^2) dpsi/dx1.
    void dihedral_force(std::vector<double> &F, int i1, int i2, int i3, int i4, double k, double psi0) {
        V3 b1 = sub(p(i2), p(i1)), b2 = sub(p(i3), p(i2)), b3 = sub(p(i4), p(i3));
        V3 m = cross(b1, b2), n = cross(b2, b3);
        double lb2 = std::sqrt(dot(b2, b2));
        double psi = std::atan2(dot(b1, n) * lb2, dot(m, n));
        double coef = 2.0 * k * (psi0 - psi);     // F = coef * dpsi/dx
        V3 g1 = scale(m, -lb2 / dot(m, m));
        V3 g4 = scale(n, lb2 / dot(n, n));
        double p12 = dot(b1, b2) / dot(b2, b2), p32 = dot(b3, b2) / dot(b2, b2);
        V3 g2 = sub(scale(g4, p32), scale(g1, 1.0 + p12));
        V3 g3 = sub(scale(g1, p12), scale(g4, 1.0 + p32));
        addf(F, i1, scale(g1, coef));
        addf(F, i2, scale(g2, coef));
        addf(F, i3, scale(g3, coef));
        addf(F, i4, scale(g4, coef));
    }
    // reduce_forces (fast.cpp:880-897): remove the WCA force of a bonded pair.
    void lj_exclusion(std::vector<double> &F, int i1, int i2, double &e_wca) {
        V3 d = pair_vec(i1, i2);
        double r2 = dot(d, d);
        double t = -lj_force(r2, type[i1], type[i2]);
        addf(F, i1, scale(d, t));
        addf(F, i2, scale(d, -t));
        e_wca -= wca_energy(r2, type[i1], type[i2]);
    }

    static double wrap_target(double psi, double psi0) {
        if (std::fabs(psi - psi0) <= kPi) return psi0;
        if (psi - psi0 > kPi) return psi0 + 2.0 * kPi;
        return psi0 - 2.0 * kPi;
    }

    // ------------------------------------------------------------ terms
    // stacking_interactions (fast.cpp:1911-2141)
    void stacking(std::vector<double> &F) {
        std::vector<char> off(n2, 0);
        std::fill(st_energy.begin(), st_energy.end(), 0.0);
        for (int k = 0; k < n3; ++k) {
            const int *b = &st3b[6 * k];
            const double *g = &st3g[6 * k];
            double dr = dist(b[1], b[4]) - g[0];
            if (dr < 10.0) {
                for (int c = 0; c < 2; ++c) {
                    int face = st3face[2 * k + c];
                    if (st_excess[face - 1] == 1) {
                        int m = (face + 1) / 2 - 2;   // s3_N + ceil(face/2) - 1, 0-based secondary
                        if (m >= 0 && m < n2) off[m] = 1;
                    }
                }
            }
            double th1 = angle(b[0], b[1], b[4]), th2 = angle(b[3], b[4], b[1]);
            double ps = dihedral(b[0], b[1], b[4], b[3]);
            double ps1 = dihedral(b[4], b[1], b[0], b[2]);
            double ps2 = dihedral(b[1], b[4], b[3], b[5]);
            double p0 = wrap_target(ps, g[3]), p10 = wrap_target(ps1, g[4]), p20 = wrap_target(ps2, g[5]);
            double q = 1.0 + 5.00 * dr * dr;
            q += 1.50 * (th1 - g[1]) * (th1 - g[1]);
            q += 1.50 * (th2 - g[2]) * (th2 - g[2]);
            q += 0.15 * (ps - p0) * (ps - p0);
            q += 0.15 * (ps1 - p10) * (ps1 - p10);
            q += 0.15 * (ps2 - p20) * (ps2 - p20);
            double e = st3E[k] / q;
            st_energy[k] = e;
            double w = -e / q;
            bond_force(F, b[1], b[4], 5.00 * w, g[0]);
            angle_force(F, b[0], b[1], b[4], 1.50 * w, g[1], true);
            angle_force(F, b[3], b[4], b[1], 1.50 * w, g[2], true);
            dihedral_force(F, b[0], b[1], b[4], b[3], 0.15 * w, p0);
            dihedral_force(F, b[4], b[1], b[0], b[2], 0.15 * w, p10);
            dihedral_force(F, b[1], b[4], b[3], b[5], 0.15 * w, p20);
            E3_ST += e;
        }
        for (int k = 0; k < n2; ++k) {
            if (off[k]) continue;
            const int *b = &st2b[7 * k];   // P_n S_n B_n P_n+1 S_n+1 B_n+1 P_n+2
            double r = dist(b[2], b[5]);
            double ps1 = dihedral(b[0], b[1], b[3], b[4]);
            double ps2 = dihedral(b[6], b[4], b[3], b[1]);
            double p10 = wrap_target(ps1, -2.58684), p20 = wrap_target(ps2, 3.07135);
            double q = 1.0 + 1.40 * (r - st2r[k]) * (r - st2r[k]);
            q += 4.00 * (ps1 - p10) * (ps1 - p10);
            q += 4.00 * (ps2 - p20) * (ps2 - p20);
            double e = st2E[k] / q;
            double w = -e / q;
            bond_force(F, b[2], b[5], 1.40 * w, st2r[k]);
            dihedral_force(F, b[0], b[1], b[3], b[4], 4.00 * w, p10);
            dihedral_force(F, b[6], b[4], b[3], b[1], 4.00 * w, p20);
            E2_ST += e;
        }
    }

    // hydrogen_bonds_interactions (fast.cpp:2145-2350)
    void hydrogen_bonds(std::vector<double> &F) {
        excess = hb_excess0;
        std::fill(hb_energy.begin(), hb_energy.end(), 0.0);
        for (int k = 0; k < nh; ++k) hb_outcome[k] = hb_status[k] == -1 ? -1 : 1;
        for (int k = 0; k < nh; ++k) {
            if (hb_status[k] == -1) continue;
            const int *b = &hbb[6 * k];
            const double *g = &hbg[6 * k];
            double r = dist(b[1], b[4]);
            if (std::fabs(r - g[0]) > 2.0) {
                hb_status[k] = 1;
                hb_outcome[k] = 0;
                for (int q = hbso[k]; q < hbso[k + 1]; ++q) excess[hbsi[q]] -= 1;
                continue;
            }
            double th1 = angle(b[0], b[1], b[4]), th2 = angle(b[3], b[4], b[1]);
            double ps = dihedral(b[0], b[1], b[4], b[3]);
            double ps1 = dihedral(b[4], b[1], b[0], b[2]);
            double ps2 = dihedral(b[1], b[4], b[3], b[5]);
            double *p0 = &hb_psi0[3 * k];
            p0[0] = wrap_target(ps, g[3]); p0[1] = wrap_target(ps1, g[4]); p0[2] = wrap_target(ps2, g[5]);
            double q = 5.00 * (r - g[0]) * (r - g[0]);
            q += 1.50 * (th1 - g[1]) * (th1 - g[1]);
            q += 1.50 * (th2 - g[2]) * (th2 - g[2]);
            q += 0.15 * (ps - p0[0]) * (ps - p0[0]);
            q += 0.15 * (ps1 - p0[1]) * (ps1 - p0[1]);
            q += 0.15 * (ps2 - p0[2]) * (ps2 - p0[2]);
            hb_energy[k] = hbE[k] * std::exp(-q);
        }
        // Valence resolution: while a site has more live HBs than its valence,
        // pick such a site at random, shuffle its live HBs, walk them keeping the
        // current pick k and moving to j with probability exp((E_j - E_k)/T),
        // then switch the pick off (fast.cpp:2239-2315).
        std::vector<int> cand_sites, bonds;
        hb_iters = 0;
        for (;;) {
            cand_sites.clear();
            for (int s = 0; s < ns; ++s)
                if (excess[s] > 0) cand_sites.push_back(s);
            int nA = (int)cand_sites.size();
            if (nA == 0) break;
            ++hb_iters;
            double r = uniform();
            int j = (int)std::ceil(r * nA);
            if (j == nA + 1) j = nA;
            if (j == 0) j = 1;
            int site = cand_sites[j - 1];
            bonds.clear();
            for (int q = shbo[site]; q < shbo[site + 1]; ++q)
                if (hb_status[shbi[q]] == 0) bonds.push_back(shbi[q]);
            int nK = (int)bonds.size();
            for (int i = 0; i < nK; ++i) {
                double u = uniform();
                int jj = (int)std::ceil(u * nK);
                if (jj == nK + 1) jj = nK;
                if (jj == 0) jj = 1;
                std::swap(bonds[i], bonds[jj - 1]);
            }
            if (nK == 0) { stream_overflow = true; break; }   // inconsistent tables; never in practice
            int k = bonds[0];
            for (int i = 1; i < nK; ++i) {
                int jb = bonds[i];
                double pr = std::exp((hb_energy[jb] - hb_energy[k]) / T);
                double u = uniform();
                if (u < pr) k = jb;
            }
            hb_status[k] = 1;
            hb_outcome[k] = 2;
            for (int q = hbso[k]; q < hbso[k + 1]; ++q) excess[hbsi[q]] -= 1;
        }
        for (int k = 0; k < nh; ++k) {
            if (hb_status[k] == -1) continue;
            if (hb_status[k] == 1) { hb_status[k] = 0; continue; }
            const int *b = &hbb[6 * k];
            const double *g = &hbg[6 * k];
            const double *p0 = &hb_psi0[3 * k];
            double w = -hb_energy[k];
            if (hbdode[k]) E3_HB -= w; else E2_HB -= w;
            bond_force(F, b[1], b[4], 5.00 * w, g[0]);
            angle_force(F, b[0], b[1], b[4], 1.50 * w, g[1], true);
            angle_force(F, b[3], b[4], b[1], 1.50 * w, g[2], true);
            dihedral_force(F, b[0], b[1], b[4], b[3], 0.15 * w, p0[0]);
            dihedral_force(F, b[4], b[1], b[0], b[2], 0.15 * w, p0[1]);
            dihedral_force(F, b[1], b[4], b[3], b[5], 0.15 * w, p0[2]);
        }
    }

    // polymer_constraints (fast.cpp:2354-2520), terms derived from the unit tables.
    void backbone(std::vector<double> &F, double &e_bond, double &e_angle, double &e_wca) {
        auto bond = [&](int i, int j, double k, double r0) {
            double r = dist(i, j);
            e_bond += k * (r - r0) * (r - r0);
            bond_force(F, i, j, k, r0);
            lj_exclusion(F, i, j, e_wca);
        };
        auto ang = [&](int i, int j, int k, double kk, double t0) {
            double th = angle(i, j, k);
            e_angle += kk * (th - t0) * (th - t0);
            angle_force(F, i, j, k, kk, t0, false);
            lj_exclusion(F, i, k, e_wca);
        };
        for (int e = 1; e <= NU; e += 2) {
            int i = u0[e - 1], j = u1[e - 1];
            double r0, a1, a2;
            switch (ukind[e - 1]) {
                case 'A': r0 = 4.8515; a1 = 1.9259; a2 = 1.7029; break;
                case 'C': r0 = 4.2738; a1 = 1.9655; a2 = 1.5803; break;
                case 'G': r0 = 4.9659; a1 = 1.9150; a2 = 1.7690; break;
                case 'U': r0 = 4.2733; a1 = 1.9663; a2 = 1.5735; break;
                default: continue;
            }
            bond(i, j, 10.0, r0);
            if (e < NU) ang(j, i, u0[e], 5.0, a1);
            if (e > 1) ang(j, i, u0[e - 2], 5.0, a2);
        }
        for (int e = 2; e <= NU; e += 2) {
            int i = u0[e - 1], j = u0[e - 2];
            bond(i, j, 64.0, 3.8157);
            if (e > 2) ang(i, j, u0[e - 3], 20.0, 1.4440);
        }
        for (int e = 3; e <= NU; e += 2) {
            int i = u0[e - 1], j = u0[e - 2], k = u0[e - 3];
            bond(i, j, 23.0, 4.6010);
            ang(i, j, k, 20.0, 1.5256);
        }
    }

    // SR_interactions_lists (fast.cpp:1783-1847)
    void short_range(std::vector<double> &F, double &e_wca, double &e_coul) {
        for (size_t q = 0; q < list_sr.size(); q += 2) {
            int i = list_sr[q], j = list_sr[q + 1];
            V3 d = pair_vec(i, j);
            double r2 = dot(d, d);
            double t = lj_force(r2, type[i], type[j]);
            e_wca += wca_energy(r2, type[i], type[j]);
            if (charged(i) && charged(j) && r2 <= kCoulCut2) {
                double r = std::sqrt(r2);
                t += pt.q_c[type[i]][type[j]] * lB / (r * r2);
                e_coul += pt.q_c[type[i]][type[j]] * lB / r;
            }
            addf(F, i, scale(d, t));
            addf(F, j, scale(d, -t));
        }
    }
    // LR_interactions_lists (fast.cpp:1851-1907)
    void long_range(std::vector<double> &F, double &e_coul) {
        for (size_t q = 0; q < list_lr.size(); q += 2) {
            int i = list_lr[q], j = list_lr[q + 1];
            V3 d = pair_vec(i, j);
            double r2 = dot(d, d);
            if (r2 > kCoulCut2) continue;
            double r = std::sqrt(r2);
            double t = pt.q_c[type[i]][type[j]] * lB / (r * r2);
            e_coul += pt.q_c[type[i]][type[j]] * lB / r;
            addf(F, i, scale(d, t));
            addf(F, j, scale(d, -t));
        }
    }

    double rg() const {
        double c[3] = {0, 0, 0};
        for (int i = 0; i < NB; ++i) for (int a = 0; a < 3; ++a) c[a] += x[3 * i + a];
        for (int a = 0; a < 3; ++a) c[a] /= NB;
        double s = 0;
        for (int i = 0; i < NB; ++i)
            for (int a = 0; a < 3; ++a) { double d = x[3 * i + a] - c[a]; s += d * d; }
        return std::sqrt(s / NB);
    }

    // deterministic_forces_MPI (fast.cpp:2524-2586) with one rank.
    void deterministic(bool lr_due) {
        std::fill(f.begin(), f.end(), 0.0);
        E2_HB = E3_HB = E2_ST = E3_ST = 0;
        stacking(f);
        hydrogen_bonds(f);
        double eb = 0, ea = 0, ew = 0, ec = 0;
        backbone(f, eb, ea, ew);
        short_range(f, ew, ec);
        if (lr_due) {
            std::fill(fl.begin(), fl.end(), 0.0);
            double el = 0;
            long_range(fl, el);
        }
        for (int i = 0; i < 3 * N; ++i) f[i] += fl[i];
    }

    // full_forces (fast.cpp:2590-2625): 1500 Box-Muller pairs, particle i (1-based)
    // takes slots i % 1000 of the x, y and z blocks.
    void noise(double noise_scale) {
        double mw[3001];
        for (int i = 1; i < 3001; i += 2) {
            double r1;
            do { r1 = uniform(); } while (r1 == 0 && !stream_overflow);
            double r2 = uniform();
            double a = std::sqrt(-2.0 * std::log(r1)), b = 2.0 * kPi * r2;
            mw[i] = a * std::cos(b);
            mw[i + 1] = a * std::sin(b);
        }
        for (int i = 0; i < N; ++i) {
            int k = (i + 1) % 1000;
            double s = sigma[type[i]] * noise_scale;
            f[3 * i] += mw[k + 1] * s;
            f[3 * i + 1] += mw[k + 1001] * s;
            f[3 * i + 2] += mw[k + 2001] * s;
        }
    }

    // move_rigid_units (fast.cpp:2660-2732): leap-frog with speed cap, recentre, wrap ions.
    void integrate() {
        for (int i = 0; i < N; ++i) {
            int t = type[i];
            double vmax = 300.0 * T;
            double vn[3], xs[3];
            for (int c = 0; c < 3; ++c) vn[c] = K1[t] * v[3 * i + c] + K2[t] * f[3 * i + c];
            double vsq = mass[t] * (vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]) / vmax;
            if (vsq > 1.0) {
                vsq = 1.0 / std::sqrt(vsq);
                for (int c = 0; c < 3; ++c) vn[c] *= vsq;
            }
            for (int c = 0; c < 3; ++c) {
                xs[c] = x[3 * i + c];
                x[3 * i + c] = 2.0 * xs[c] - xo[3 * i + c] + (vn[c] - v[3 * i + c]) * h;
                xo[3 * i + c] = xs[c];
                v[3 * i + c] = vn[c];
            }
        }
        double s[3];
        for (int c = 0; c < 3; ++c) s[c] = hL[c] - x[3 * kh + c];
        for (int i = 0; i < N; ++i)
            for (int c = 0; c < 3; ++c) { x[3 * i + c] += s[c]; xo[3 * i + c] += s[c]; }
        for (int i = NB; i < N; ++i)
            for (int c = 0; c < 3; ++c) {
                if (x[3 * i + c] < 0) { x[3 * i + c] += L[c]; xo[3 * i + c] += L[c]; }
                if (x[3 * i + c] > L[c]) { x[3 * i + c] -= L[c]; xo[3 * i + c] -= L[c]; }
            }
    }

    // check_shifts (fast.cpp:1741-1776)
    void check_shifts() {
        double m = 0;
        for (int i = 0; i < N; ++i) {
            double a = x[3 * i] - xo[3 * i], b = x[3 * i + 1] - xo[3 * i + 1], c = x[3 * i + 2] - xo[3 * i + 2];
            double r2 = a * a + b * b + c * c;
            if (r2 > m) m = r2;
        }
        m = std::sqrt(m);
        progress += m + m;
        if (progress > skin) {
            build_lists_cells();
            progress = 0;
            ++rebuilds;
        }
    }
};

extern "C" {

oracle_ctx *oracle_create(const tis_system *s, double r_sr, double r_lr, double skin, int lr_every) {
    if (!s || s->n_particles <= 0) return nullptr;
    oracle_ctx *o = new (std::nothrow) oracle_ctx();
    if (!o) return nullptr;
    o->N = s->n_particles; o->NB = s->n_beads; o->NU = s->n_units;
    o->n3 = s->n_st3; o->n2 = s->n_st2; o->nh = s->n_hb; o->ns = s->n_hb_sites;
    for (int c = 0; c < 3; ++c) { o->L[c] = s->box[c]; o->hL[c] = 0.5 * s->box[c]; }
    o->T = s->T; o->lB = s->lB; o->h = s->h; o->kh = s->recentre;
    int N = o->N;
    o->type.assign(s->type, s->type + N);
    o->cls.assign(s->cls, s->cls + N);
    o->u0.assign(s->unit_bead0, s->unit_bead0 + o->NU);
    o->u1.assign(s->unit_bead1, s->unit_bead1 + o->NU);
    o->ukind.assign(s->unit_kind, s->unit_kind + o->NU);
    // read_maxi_key derived tables (fast.cpp:2895-2944)
    for (int a = 0; a < 9; ++a) {
        double visc = 0.555 * s->type_R[a];
        o->mass[a] = s->type_mass[a];
        o->K1[a] = std::exp(-visc / s->type_mass[a] * s->h);
        o->K2[a] = (1.0 - o->K1[a]) / visc;
        o->sigma[a] = std::sqrt(2.0 * visc * s->T / s->h);
    }
    for (int a = 0; a < 9; ++a)
        for (int b = 0; b < 9; ++b) {
            int lo = std::min(a, b), hi = std::max(a, b);
            double d = (lo <= 5 && hi <= 5) ? 3.2 : s->type_R[lo] + s->type_R[hi];
            o->pt.d_lj[a][b] = d;
            o->pt.d2_lj[a][b] = d * d;
            double dm = d - 0.4406141861 * kSig;
            o->pt.d2_min[a][b] = dm * dm;
            o->pt.e_lj[a][b] = std::sqrt(s->type_eps[lo] * s->type_eps[hi]);
            o->pt.f_lj[a][b] = 12.0 * o->pt.e_lj[a][b] / (kSig * kSig);
            o->pt.q_c[a][b] = s->type_q[lo] * s->type_q[hi];
        }
    o->st3b.assign(s->st3_beads, s->st3_beads + 6 * o->n3);
    o->st3g.assign(s->st3_geo, s->st3_geo + 6 * o->n3);
    o->st3E.assign(s->st3_E, s->st3_E + o->n3);
    o->st3face.assign(s->st3_face, s->st3_face + 2 * o->n3);
    o->st2b.assign(s->st2_beads, s->st2_beads + 7 * o->n2);
    o->st2r.assign(s->st2_r0, s->st2_r0 + o->n2);
    o->st2E.assign(s->st2_E, s->st2_E + o->n2);
    o->st_excess.assign(s->st_excess, s->st_excess + o->NU);
    o->hbb.assign(s->hb_beads, s->hb_beads + 6 * o->nh);
    o->hbg.assign(s->hb_geo, s->hb_geo + 6 * o->nh);
    o->hbE.assign(s->hb_E, s->hb_E + o->nh);
    o->hbdode.assign(s->hb_dode, s->hb_dode + o->nh);
    o->hbso.assign(s->hb_site_off, s->hb_site_off + o->nh + 1);
    o->hbsi.assign(s->hb_site_idx, s->hb_site_idx + o->hbso[o->nh]);
    o->shbo.assign(s->site_hb_off, s->site_hb_off + o->ns + 1);
    o->shbi.assign(s->site_hb_idx, s->site_hb_idx + o->shbo[o->ns]);
    o->valence.assign(s->site_valence, s->site_valence + o->ns);
    o->r_sr2 = r_sr * r_sr;
    o->r_lr2 = r_lr * r_lr;
    o->r_cell = r_lr;
    o->skin = skin;
    o->lr_every = lr_every;
    o->hb_status.assign(o->nh, -1);
    o->hb_excess0.assign(o->ns, 0);
    o->hb_outcome.assign(o->nh, -1);
    o->hb_energy.assign(o->nh, 0.0);
    o->hb_psi0.assign(3 * (size_t)o->nh, 0.0);
    o->st_energy.assign(o->n3, 0.0);
    o->x.assign(3 * (size_t)N, 0.0);
    o->xo.assign(3 * (size_t)N, 0.0);
    o->v.assign(3 * (size_t)N, 0.0);
    o->f.assign(3 * (size_t)N, 0.0);
    o->fl.assign(3 * (size_t)N, 0.0);
    return o;
}

void oracle_destroy(oracle_ctx *o) { delete o; }

int oracle_set_state(oracle_ctx *o, const double *pos, const double *pos_old, const double *vel,
                     long long klok, int brute) {
    if (!o || !pos) return TIS_EINVAL;
    size_t n = 3 * (size_t)o->N;
    std::memcpy(o->x.data(), pos, n * sizeof(double));
    if (pos_old) std::memcpy(o->xo.data(), pos_old, n * sizeof(double));
    if (vel) std::memcpy(o->v.data(), vel, n * sizeof(double));
    o->klok = klok;
    o->progress = 0;
    if (brute) o->build_lists_brute(); else o->build_lists_cells();
    return TIS_OK;
}

int oracle_get_state(const oracle_ctx *o, double *pos, double *pos_old, double *vel) {
    if (!o) return TIS_EINVAL;
    size_t n = 3 * (size_t)o->N * sizeof(double);
    if (pos) std::memcpy(pos, o->x.data(), n);
    if (pos_old) std::memcpy(pos_old, o->xo.data(), n);
    if (vel) std::memcpy(vel, o->v.data(), n);
    return TIS_OK;
}

int oracle_set_stream(oracle_ctx *o, const int32_t *raw, long long n) {
    if (!o || (n && !raw)) return TIS_EINVAL;
    o->stream.assign(raw, raw + n);
    o->cursor = 0;
    o->stream_overflow = false;
    return TIS_OK;
}

long long oracle_stream_cursor(const oracle_ctx *o) { return o ? o->cursor : -1; }

int oracle_stream_glibc_more(oracle_ctx *o, long long n) {
    if (!o || n < 0) return TIS_EINVAL;
    o->stream.erase(o->stream.begin(), o->stream.begin() + std::min<long long>(o->cursor, (long long)o->stream.size()));
    for (long long q = 0; q < n; ++q) o->stream.push_back(std::rand());
    o->cursor = 0;
    o->stream_overflow = false;
    return TIS_OK;
}

int oracle_set_stream_glibc(oracle_ctx *o, unsigned seed, long long n) {
    if (!o || n < 0) return TIS_EINVAL;
    o->stream.resize(n);
    std::srand(seed);
    for (long long q = 0; q < n; ++q) o->stream[q] = std::rand();
    o->cursor = 0;
    o->stream_overflow = false;
    return TIS_OK;
}

int oracle_eval(oracle_ctx *o, unsigned mask, double *f_out, double *e_out) {
    if (!o) return TIS_EINVAL;
    std::vector<double> F(3 * (size_t)o->N, 0.0);
    double e[OE_N] = {0};
    o->E2_HB = o->E3_HB = o->E2_ST = o->E3_ST = 0;
    if (mask & TIS_TERM_STACK) o->stacking(F);
    if (mask & TIS_TERM_HB) o->hydrogen_bonds(F);
    if (mask & TIS_TERM_BACKBONE) o->backbone(F, e[OE_BOND], e[OE_ANGLE], e[OE_WCA]);
    if (mask & TIS_TERM_SR) o->short_range(F, e[OE_WCA], e[OE_COUL_SR]);
    if (mask & TIS_TERM_LR) o->long_range(F, e[OE_COUL_LR]);
    e[OE_E2_HB] = o->E2_HB; e[OE_E3_HB] = o->E3_HB;
    e[OE_E2_STACK] = o->E2_ST; e[OE_E3_STACK] = o->E3_ST;
    e[OE_RG] = o->rg();
    if (f_out) std::memcpy(f_out, F.data(), F.size() * sizeof(double));
    if (e_out) std::memcpy(e_out, e, sizeof e);
    return o->stream_overflow ? TIS_EOVERFLOW : TIS_OK;
}

int oracle_step(oracle_ctx *o, long long n, double noise_scale) {
    if (!o) return TIS_EINVAL;
    for (long long s = 0; s < n; ++s) {
        bool lr_due = (o->klok % o->lr_every) == 1 || o->progress == 0;
        o->deterministic(lr_due);
        o->noise(noise_scale);
        o->integrate();
        o->check_shifts();
        ++o->klok;
        if (o->stream_overflow) return TIS_EOVERFLOW;
    }
    return TIS_OK;
}

long long oracle_pairs(const oracle_ctx *o, int which, int *out, long long cap) {
    if (!o) return -1;
    const std::vector<int> &l = which ? o->list_lr : o->list_sr;
    long long np = (long long)l.size() / 2;
    std::vector<std::pair<int, int>> v(np);
    for (long long q = 0; q < np; ++q) {
        int a = l[2 * q], b = l[2 * q + 1];
        v[q] = {std::min(a, b), std::max(a, b)};
    }
    std::sort(v.begin(), v.end());
    if (out) {
        long long m = std::min(np, cap);
        for (long long q = 0; q < m; ++q) { out[2 * q] = v[q].first; out[2 * q + 1] = v[q].second; }
    }
    return np;
}

int oracle_hb_status(const oracle_ctx *o, int *out) {
    if (!o || !out) return TIS_EINVAL;
    std::copy(o->hb_status.begin(), o->hb_status.end(), out);
    return TIS_OK;
}

long long oracle_rebuilds(const oracle_ctx *o) { return o ? o->rebuilds : -1; }
double oracle_progress(const oracle_ctx *o) { return o ? o->progress : -1; }
long long oracle_klok(const oracle_ctx *o) { return o ? o->klok : -1; }
int oracle_stack_energies(const oracle_ctx *o, double *out) {
    if (!o || !out) return TIS_EINVAL;
    std::copy(o->st_energy.begin(), o->st_energy.end(), out);
    return TIS_OK;
}
long long oracle_hb_iterations(const oracle_ctx *o) { return o ? o->hb_iters : -1; }
int oracle_hb_outcome(const oracle_ctx *o, int *out) {
    if (!o || !out) return TIS_EINVAL;
    std::copy(o->hb_outcome.begin(), o->hb_outcome.end(), out);
    return TIS_OK;
}

}  // extern "C"
