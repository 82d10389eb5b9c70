// Host set-up of the 16SCD TIS system: the part of the reference that runs
// once, before the Langevin loop.  It reads the reference's own input files and
// reproduces its tables, geometry and glibc rand() consumption, so that the
// state handed to the GPU is the one the reference would start from.
//
// Reference routines restated here (fast.cpp = src/16SCD_Mg_Cl_K_COPY_fast.cpp):
//   read_maxi_key                fast.cpp:2883-2945
//   read_PDB                     fast.cpp:3137-3632 (+ generate_atom_velocity :509-566)
//   return_HB_JNDX               fast.cpp:3636-3667
//   initialize_HB_atoms          fast.cpp:3671-3729
//   initialize_unprocessed_bonds fast.cpp:3733-4119
//   initialize_unprocessed_stacks fast.cpp:4232-4492
//   set_box                      fast.cpp:4542-5051
//   add_crowder (single-site ions) fast.cpp:5055-5403
//   ion counts / init order      fast.cpp:6105-6177
//   backbone term order          fast.cpp:2354-2520 (polymer_constraints)
//
// Internally the tables are 1-based like the reference's (slot 0 unused), so
// the unit arithmetic (2*RES-1, atom_key[i+1][0], ...) reads the same; they are
// converted to the 0-based tis_system at the end.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/tis_gpu.h"
#include "glibc_rand.h"
#include "host_util.h"

namespace tis_host {
namespace {

constexpr double kPi = 3.14159265358979323846264338328;
constexpr double kSigmaLJ = 1.5852;   // shift length of the WCA potential, fast.cpp:801-821
constexpr double kHbUnit = 0.73753;   // hb_E per atom-level hydrogen bond, fast.cpp:3881

// Whitespace token reader with the reference's "read 9 tokens, drop the rest of
// the line" PDB convention (fast.cpp:3163-3282: fscanf("%s") x9 then fgets).
class TokenFile {
  public:
    explicit TokenFile(const std::string &path) {
        std::ifstream in(path, std::ios::binary);
        if (!in) throw IoError("cannot open " + path);
        std::ostringstream ss;
        ss << in.rdbuf();
        buf_ = ss.str();
    }
    bool next(std::string &tok) {
        while (p_ < buf_.size() && std::isspace((unsigned char)buf_[p_])) ++p_;
        if (p_ >= buf_.size()) return false;
        size_t s = p_;
        while (p_ < buf_.size() && !std::isspace((unsigned char)buf_[p_])) ++p_;
        tok.assign(buf_, s, p_ - s);
        return true;
    }
    void skip_line() {
        while (p_ < buf_.size() && buf_[p_] != '\n') ++p_;
        if (p_ < buf_.size()) ++p_;
    }
    void rewind() { p_ = 0; }

  private:
    std::string buf_;
    size_t p_ = 0;
};

// One PDB record as the reference sees it: tokens 1..9 of a line.
struct PdbRecord {
    std::string name, resname, resnum;
    std::string xs, ys, zs;
};

std::vector<PdbRecord> read_pdb_records(const std::string &path) {
    TokenFile tf(path);
    std::vector<PdbRecord> recs;
    std::string t[9];
    for (;;) {
        int got = 0;
        while (got < 9 && tf.next(t[got])) ++got;
        if (got == 0) break;
        if (got < 9) throw IoError("truncated PDB record in " + path);
        tf.skip_line();
        recs.push_back({t[2], t[3], t[5], t[6], t[7], t[8]});
    }
    return recs;
}

struct Vec3 { double x, y, z; };

// --- geometry primitives on the host (same formulas as the set-up in the
// reference: CC_distance :657-670, valence_angle :986-1027, dihedral_angle
// :1255-1302; no periodic image, the chain is never wrapped).
double dist(const Vec3 &a, const Vec3 &b) {
    double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}
double angle(const Vec3 &a, const Vec3 &b, const Vec3 &c) {
    double v1[3] = {a.x - b.x, a.y - b.y, a.z - b.z};
    double v2[3] = {c.x - b.x, c.y - b.y, c.z - b.z};
    double n1 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
    double n2 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];
    double n12 = std::sqrt(n1 * n2);
    double c0 = (v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2]) / n12;
    return std::acos(c0);
}
void cross(const double *a, const double *b, double *r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
double dihedral(const Vec3 &p1, const Vec3 &p2, const Vec3 &p3, const Vec3 &p4) {
    double v1[3] = {p2.x - p1.x, p2.y - p1.y, p2.z - p1.z};
    double v2[3] = {p3.x - p2.x, p3.y - p2.y, p3.z - p2.z};
    double v3[3] = {p4.x - p3.x, p4.y - p3.y, p4.z - p3.z};
    double m[3], n[3];
    cross(v1, v2, m);
    cross(v2, v3, n);
    double b = std::sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
    double c0 = m[0] * n[0] + m[1] * n[1] + m[2] * n[2];
    double s0 = (v1[0] * n[0] + v1[1] * n[1] + v1[2] * n[2]) * b;
    return std::atan2(s0, c0);
}

// Secondary-stack parameters by (5' base, 3' base): r0 and the free energy
// E = -dG + ss_D / c + s * (T - Tm)  (fast.cpp:4399-4429).
struct SecStackParam { char a, b; double r0, dG, c, s, Tm; };
const SecStackParam kSecStack[16] = {
    {'A', 'A', 4.1806530, 5.194, 0.7093838769, -0.319, 0.594},
    {'A', 'C', 3.8260185, 5.146, 0.7185955600, -0.319, 0.594},
    {'A', 'G', 4.4255305, 5.977, 0.6968019829, 5.301, 0.678},
    {'A', 'U', 3.8260185, 5.146, 0.7185955600, -0.319, 0.594},
    {'C', 'A', 4.7010580, 5.163, 0.6847830171, -0.319, 0.594},
    {'C', 'C', 4.2500910, 4.873, 0.6991615586, -1.567, 0.568},
    {'C', 'G', 4.9790760, 5.482, 0.6816268897, 0.774, 0.627},
    {'C', 'U', 4.2273615, 4.873, 0.6832570771, -1.567, 0.568},
    {'G', 'A', 4.0128560, 5.948, 0.6903176657, 5.301, 0.678},
    {'G', 'C', 3.6784360, 5.927, 0.7042060343, 4.370, 0.682},
    {'G', 'G', 4.2427250, 6.416, 0.6971421514, 7.346, 0.728},
    {'G', 'U', 3.6616930, 5.836, 0.6984426543, 2.924, 0.672},
    {'U', 'A', 4.7010580, 5.163, 0.6847830171, -0.319, 0.594},
    {'U', 'C', 4.2679180, 4.880, 0.6758595771, -1.567, 0.568},
    {'U', 'G', 4.9977560, 5.886, 0.7025528229, 2.924, 0.672},
    {'U', 'U', 4.2453650, 4.267, 0.6686014771, -3.563, 0.500},
};

// Canonical base-pair HB geometries (r, theta1, theta2, psi, psi1, psi2) that
// replace the PDB-measured ones for CANxy codes (fast.cpp:3967-3985).
struct CanonicalHb { const char *code; double g[6]; };
const CanonicalHb kCanonical[10] = {
    {"CAN11", {5.8815, 2.7283, 2.5117, 1.2559, 0.9545, 1.1747}},
    {"CAN12", {5.8815, 2.5117, 2.7283, 1.2559, 1.1747, 0.9545}},
    {"CAN21", {5.6550, 2.4837, 2.8230, 1.3902, 1.2174, 0.7619}},
    {"CAN22", {5.6550, 2.8230, 2.4837, 1.3902, 0.7619, 1.2174}},
    {"CAN31", {5.9234, 2.9051, 2.0544, 2.1469, -0.5366, 1.4288}},
    {"CAN32", {5.9234, 2.0544, 2.9051, 2.1469, 1.4288, -0.5366}},
    {"CAN41", {5.3737, 2.8295, 2.0314, 0.6063, 1.0893, 1.4835}},
    {"CAN42", {5.3737, 2.0314, 2.8295, 0.6063, 1.4835, 1.0893}},
    {"CAN51", {6.4922, 2.0579, 1.3569, 3.1147, -1.4796, 1.5278}},
    {"CAN52", {6.4922, 1.3569, 2.0579, 3.1147, 1.5278, -1.4796}},
};

// 16SCD residues (PDB numbering) whose 5' stacking face, unit 2*n-2 of
// nucleotide n, has ST_EXCESS lowered by 2 instead of 1: the list the
// reference hard-codes in fast.cpp:4458-4491.
const int k16scdDoubleFace[15] = {583, 587, 658, 687, 700, 701, 721, 727, 760, 782, 802, 814, 881, 892, 898};

struct Builder {
    tis_build_args args{};
    GlibcRand rng;
    // per type (1..9)
    double mass[10]{}, rlj[10]{}, elj[10]{}, qc[10]{};
    // particles (1-based)
    std::vector<double> x{0}, y{0}, z{0}, xo{0}, yo{0}, zo{0}, vx{0}, vy{0}, vz{0};
    std::vector<int> part{0}, maxi{0};
    // units (1-based)
    std::vector<char> kind{0};
    std::vector<int> key0{0}, key1{0};
    int n_units = 0, n_beads = 0;
    // HB sites (1-based), per unit first site index
    std::vector<std::vector<int>> unit_sites{{}};
    std::vector<int> valence{0};
    // HB candidates (1-based)
    struct Hb { int b[6]; double g[6]; double E; int dode; std::vector<int> sites; };
    std::vector<Hb> hb{Hb{}};
    std::vector<std::vector<int>> site_hbs{{}};
    // stacks (1-based)
    struct St3 { int unit_i, unit_j; int b[6]; double g[6]; double E; int face[2]; };
    std::vector<St3> st3{St3{}};
    struct St2 { int unit; double r0, E; };
    std::vector<St2> st2{St2{}};
    std::vector<int> st_excess;
    double box = 350.0, volume = 0, half[3]{};
    int n_ion[3]{};
    int kh = 500;                 // recentring bead (1-based), fast.cpp:6163
    // PDB residue number -> nucleotide index (1-based) in chain order.  The
    // reference hard-codes 16SCD's: res - 561, or res - 567 from 842 on
    // (fast.cpp:3783-3795, 4295-4305), so a number inside the 842-847 gap
    // takes the offset of the segment after the gap.  The same rule in general:
    // a listed number maps to its PDB position, a number missing from the PDB
    // to the offset of the next number the PDB lists (of the last one past
    // the end), and the result must be a nucleotide of the chain.
    std::map<int, int> res_index;

    int residue(int res) const {
        if (res_index.empty()) throw IoError("no residues read from the PDB");
        auto it = res_index.lower_bound(res);
        if (it == res_index.end()) it = std::prev(res_index.end());
        int idx = res - (it->first - it->second);
        if (idx < 1 || idx > (int)res_index.size())
            throw IoError("residue " + std::to_string(res) + " lies outside the chain");
        return idx;
    }

    Vec3 P(int k) const { return {x[k], y[k], z[k]}; }

    void read_maxi_key(const std::string &path) {
        TokenFile tf(path);
        std::string t;
        for (int a = 1; a <= TIS_NTYPES; ++a) {
            std::string f[5];
            for (int c = 0; c < 5; ++c)
                if (!tf.next(f[c])) throw IoError("short MAXIKEY file " + path);
            mass[a] = std::strtod(f[1].c_str(), nullptr);
            rlj[a] = std::strtod(f[2].c_str(), nullptr);
            elj[a] = std::strtod(f[3].c_str(), nullptr);
            qc[a] = std::strtod(f[4].c_str(), nullptr);
        }
    }

    static int pair_index(int a, int b) {  // 1-based triangular Jn12 (fast.cpp:869-871)
        if (a > b) std::swap(a, b);
        return (a - 1) * TIS_NTYPES - (a - 1) * a / 2 + b;
    }
    double d_lj(int a, int b) const {
        return (a <= 6 && b <= 6) ? 3.2 : rlj[a] + rlj[b];
    }
    double d2_overlap(int a, int b) const {
        double d = d_lj(a, b) - 0.2 * kSigmaLJ;
        return d * d;
    }

    double uniform() { return (double)rng.next() / GlibcRand::kRandMax; }

    // generate_atom_velocity (fast.cpp:509-566): three Box-Muller cosines.
    void draw_velocity(int k) {
        double m = mass[maxi[k]];
        double *v[3] = {&vx[k], &vy[k], &vz[k]};
        for (int c = 0; c < 3; ++c) {
            double a1;
            do { a1 = uniform(); } while (a1 == 0);
            double a2 = uniform();
            double t1 = std::sqrt(args.T / m) * std::sqrt(-2.0 * std::log(a1));
            double t2 = 2.0 * kPi * a2;
            *v[c] = t1 * std::cos(t2);
        }
        xo[k] = x[k] - vx[k] * 0.05;
        yo[k] = y[k] - vy[k] * 0.05;
        zo[k] = z[k] - vz[k] * 0.05;
    }

    int new_particle(int cls, int type) {
        x.push_back(0); y.push_back(0); z.push_back(0);
        xo.push_back(0); yo.push_back(0); zo.push_back(0);
        vx.push_back(0); vy.push_back(0); vz.push_back(0);
        part.push_back(cls);
        maxi.push_back(type);
        return (int)x.size() - 1;
    }

    void read_pdb(const std::string &path) {
        std::vector<PdbRecord> recs = read_pdb_records(path);
        // Pass 1: units and their kinds.
        for (const PdbRecord &r : recs) {
            if (r.name == "P") {
                kind.push_back('P');
            } else if (r.name == "O5'") {
                char k = 0;
                if (r.resname == "A" || r.resname == "G" || r.resname == "C" || r.resname == "U")
                    k = r.resname[0];
                if (!k) throw IoError("unknown nucleotide " + r.resname);
                kind.push_back(k);
                int nt = (int)res_index.size() + 1;
                if (!res_index.emplace(std::atoi(r.resnum.c_str()), nt).second)
                    throw IoError("residue " + r.resnum + " appears twice in the PDB");
            }
        }
        n_units = (int)kind.size() - 1;
        key0.assign(n_units + 1, 0);
        key1.assign(n_units + 1, 0);
        // Pass 2: bead centroids of the non-hydrogen atoms in token order.
        int unit = 0, cur = 0, count = 0;
        bool skip_h = false;
        auto close_bead = [&]() {
            if (cur && count) { x[cur] /= count; y[cur] /= count; z[cur] /= count; }
        };
        for (const PdbRecord &r : recs) {
            const std::string &nm = r.name;
            bool is_base_start = unit > 0 &&
                ((nm == "N9" && (kind[unit] == 'A' || kind[unit] == 'G')) ||
                 (nm == "N1" && (kind[unit] == 'C' || kind[unit] == 'U')));
            if (nm == "P") {
                skip_h = false;
                close_bead();
                ++unit;
                cur = new_particle(0, 1);
                key0[unit] = cur;
                key1[unit] = -1;
                count = 1;
            } else if (nm == "O5'") {
                close_bead();
                ++unit;
                cur = new_particle(1, 2);
                key0[unit] = cur;
                count = 1;
            } else if (is_base_start) {
                close_bead();
                int t = kind[unit] == 'A' ? 3 : kind[unit] == 'G' ? 4 : kind[unit] == 'C' ? 5 : 6;
                cur = new_particle(2, t);
                key1[unit] = cur;
                count = 1;
            } else if (nm[0] != 'H') {
                ++count;
            } else {
                skip_h = true;
            }
            if (!skip_h && cur) {
                x[cur] += std::atof(r.xs.c_str());
                y[cur] += std::atof(r.ys.c_str());
                z[cur] += std::atof(r.zs.c_str());
            }
        }
        close_bead();
        n_beads = (int)x.size() - 1;
        // Velocities unit by unit, S before B (fast.cpp:3625-3627).
        for (int u = 1; u <= n_units; ++u) {
            draw_velocity(key0[u]);
            if (kind[u] != 'P') draw_velocity(key1[u]);
        }
    }

    // return_HB_JNDX (fast.cpp:3636-3667): site slot of an atom within its unit.
    static int site_slot(const std::string &atom, char nt) {
        if (atom == "N1") return 2;
        if (atom == "N2") return 3;
        if (atom == "N3") return nt == 'G' ? 4 : 3;
        if (atom == "N4") return 4;
        if (atom == "N6") return 4;
        if (atom == "N7") return nt == 'A' ? 5 : 6;
        if (atom == "O2") return 2;
        if (atom == "O4") return 4;
        if (atom == "O6") return 5;
        if (atom == "O2'") return 0;
        if (atom == "O4'") return 1;
        if (atom == "OP1") return 0;
        if (atom == "OP2") return 1;
        throw IoError("unknown HB atom " + atom);
    }

    // initialize_HB_atoms (fast.cpp:3671-3729): sites per unit and valences.
    void init_hb_sites() {
        unit_sites.assign(n_units + 1, {});
        for (int u = 1; u <= n_units; ++u) {
            int ns = 0, v2a = -1, v2b = -1, v2c = -1;
            switch (kind[u]) {
                case 'A': ns = 5; v2a = 4; v2b = 0; v2c = -1; break;
                case 'C': ns = 4; v2a = 4; v2b = 0; v2c = 2; break;
                case 'G': ns = 6; v2a = 3; v2b = 0; v2c = 5; break;
                case 'U': ns = 4; v2a = -1; v2b = 0; v2c = 2; break;
                default: ns = 1; break;
            }
            for (int j = 0; j <= ns; ++j) {
                valence.push_back((j == v2a || j == v2b || j == v2c) ? 2 : 1);
                unit_sites[u].push_back((int)valence.size() - 1);
            }
        }
        site_hbs.assign(valence.size(), {});
    }

    // Beads of unit u; an interaction that needs a unit past either chain end
    // (the 5' phosphate of nucleotide 1, the 3' units of the last one) is an
    // input error.
    int bead0(int u) const {
        if (u < 1 || u > n_units)
            throw IoError("an HB or stack needs unit " + std::to_string(u) + ", outside the chain's 1.." +
                          std::to_string(n_units));
        return key0[u];
    }
    int bead1(int u) const { return bead0(u), key1[u]; }

    // Bead triple (k?0, k?1, k?2) of an HB atom by its chemical class (fast.cpp:3809-3863).
    void hb_triple(const std::string &atom, int res, int out[3]) const {
        if (atom == "OP1" || atom == "OP2") {
            out[0] = bead0(2 * res - 1); out[1] = bead0(2 * res - 2); out[2] = bead0(2 * res);
        } else if (atom == "O2'" || atom == "O4'") {
            out[0] = bead0(2 * res); out[1] = bead0(2 * res - 1); out[2] = bead0(2 * res + 1);
        } else {
            out[0] = bead0(2 * res - 1); out[1] = bead1(2 * res - 1); out[2] = bead0(2 * res);
        }
    }
    int hb_site(const std::string &atom, int res) const {
        int u = (atom == "OP1" || atom == "OP2") ? 2 * res - 2 : 2 * res - 1;
        return unit_sites.at(u).at(site_slot(atom, kind[u]));
    }

    // initialize_unprocessed_bonds (fast.cpp:3733-4119).
    void read_hb_candidates(const std::string &path) {
        TokenFile tf(path);
        std::string code, r1s, r2s, a1, a2;
        while (tf.next(code)) {
            if (!tf.next(r1s) || !tf.next(r2s) || !tf.next(a1)) throw IoError("short HB line in " + path);
            int res1 = residue(std::atoi(r1s.c_str()));
            int res2 = residue(std::atoi(r2s.c_str()));
            int first_of_line = (int)hb.size();   // merging only inside one line
            const CanonicalHb *can = nullptr;
            for (const CanonicalHb &c : kCanonical)
                if (code == c.code) can = &c;
            bool tertiary = code == "NON00" || code == "APLAT";
            do {
                if (!tf.next(a2)) throw IoError("unterminated HB line in " + path);
                int t1[3], t2[3];
                hb_triple(a1, res1, t1);
                hb_triple(a2, res2, t2);
                int b[6] = {t1[0], t1[1], t1[2], t2[0], t2[1], t2[2]};
                int k = -1;
                for (int q = first_of_line; q < (int)hb.size(); ++q)
                    if (std::memcmp(hb[q].b, b, sizeof b) == 0) { k = q; break; }
                if (k >= 0) {
                    hb[k].E -= args.hs_D * kHbUnit;
                } else if ((code == "CAN51" && a2 == "O2'") || (code == "CAN52" && a1 == "O2'")) {
                    k = (int)hb.size() - 1;   // the pair joins the last HB without adding energy
                } else {
                    Hb h{};
                    std::memcpy(h.b, b, sizeof b);
                    h.g[0] = dist(P(b[1]), P(b[4]));
                    h.g[1] = angle(P(b[0]), P(b[1]), P(b[4]));
                    h.g[2] = angle(P(b[3]), P(b[4]), P(b[1]));
                    h.g[3] = dihedral(P(b[0]), P(b[1]), P(b[4]), P(b[3]));
                    h.g[4] = dihedral(P(b[4]), P(b[1]), P(b[0]), P(b[2]));
                    h.g[5] = dihedral(P(b[1]), P(b[4]), P(b[3]), P(b[5]));
                    h.E = -args.hs_D * kHbUnit;
                    if (can)
                        for (int c = 0; c < 6; ++c) h.g[c] = can->g[c];
                    h.dode = tertiary ? 1 : 0;
                    hb.push_back(h);
                    k = (int)hb.size() - 1;
                }
                int s1 = hb_site(a1, res1), s2 = hb_site(a2, res2);
                hb[k].sites.push_back(s1);
                hb[k].sites.push_back(s2);
                site_hbs[s1].push_back(k);
                site_hbs[s2].push_back(k);
                if (!tf.next(a1)) throw IoError("unterminated HB line in " + path);
            } while (a1 != "TER");
        }
    }

    // initialize_unprocessed_stacks (fast.cpp:4232-4492).
    void read_stacks(const std::string &path, const std::vector<int> &double_face_res) {
        TokenFile tf(path);
        st_excess.assign(n_units + 1, 0);
        std::string s1, s2;
        while (tf.next(s1)) {
            if (!tf.next(s2)) throw IoError("odd token count in " + path);
            int res1 = residue(std::atoi(s1.c_str() + 1));
            int res2 = residue(std::atoi(s2.c_str() + 1));
            St3 s{};
            s.unit_i = 2 * res1 - 1;
            s.unit_j = 2 * res2 - 1;
            int i = s.unit_i, j = s.unit_j;
            int b[6] = {bead0(i), bead1(i), bead0(i + 1), bead0(j), bead1(j), bead0(j + 1)};
            std::memcpy(s.b, b, sizeof b);
            s.g[0] = dist(P(b[1]), P(b[4]));
            s.g[1] = angle(P(b[0]), P(b[1]), P(b[4]));
            s.g[2] = angle(P(b[3]), P(b[4]), P(b[1]));
            s.g[3] = dihedral(P(b[0]), P(b[1]), P(b[4]), P(b[3]));
            s.g[4] = dihedral(P(b[4]), P(b[1]), P(b[0]), P(b[2]));
            s.g[5] = dihedral(P(b[1]), P(b[4]), P(b[3]), P(b[5]));
            s.E = -args.st_D;
            s.face[0] = s1[0] == '-' ? i - 1 : i;
            s.face[1] = s2[0] == '-' ? j - 1 : j;
            st_excess.at(s.face[0]) += 1;
            st_excess.at(s.face[1]) += 1;
            st3.push_back(s);
        }
        for (int i = 3; i < n_units - 2; i += 2) {
            const SecStackParam *p = nullptr;
            for (const SecStackParam &q : kSecStack)
                if (q.a == kind[i] && q.b == kind[i + 2]) p = &q;
            if (!p) throw IoError("no secondary-stack parameters for unit pair");
            St2 s{};
            s.unit = i;
            s.r0 = p->r0;
            s.E = -p->dG + args.ss_D / p->c + p->s * (args.T - p->Tm);
            st2.push_back(s);
            st_excess[i] += 1;
            st_excess[i + 1] += 1;
        }
        std::vector<int> drop(n_units + 1, 1);
        for (int r : double_face_res) {
            int f = 2 * residue(r) - 2;
            if (f >= 1) drop[f] = 2;
        }
        for (int f = 1; f <= n_units; ++f) st_excess[f] -= drop[f];
    }

    // set_box (fast.cpp:4542-5051): rotate the chain onto its principal axes
    // (smallest moment first) and move bead kh to the box centre.
    void set_box(int kh) {
        int n = n_beads;
        double a0 = 0, a1 = 0, a2 = 0;
        for (int i = 1; i <= n; ++i) { a0 += x[i]; a1 += y[i]; a2 += z[i]; }
        a0 /= n; a1 /= n; a2 /= n;
        double A = 0, B = 0, C = 0, F = 0, G = 0, H = 0;
        for (int i = 1; i <= n; ++i) {
            double dx = x[i] - a0, dy = y[i] - a1, dz = z[i] - a2;
            A += dy * dy + dz * dz;
            B += dx * dx + dz * dz;
            C += dx * dx + dy * dy;
            F += dy * dz;
            G += dx * dz;
            H += dx * dy;
        }
        double e1[3], e2[3], e3[3];
        principal_axes(A, B, C, F, G, H, e1, e2, e3);
        for (int i = 1; i <= n; ++i) {
            double px = x[i] * e1[0] + y[i] * e1[1] + z[i] * e1[2];
            double py = x[i] * e2[0] + y[i] * e2[1] + z[i] * e2[2];
            double pz = x[i] * e3[0] + y[i] * e3[1] + z[i] * e3[2];
            x[i] = px; y[i] = py; z[i] = pz;
        }
        volume = box * box * box;
        half[0] = half[1] = half[2] = 0.5 * box;
        double sx = half[0] - x[kh], sy = half[1] - y[kh], sz = half[2] - z[kh];
        for (int i = 1; i <= n; ++i) {
            x[i] += sx; y[i] += sy; z[i] += sz;
            double ox = xo[i] * e1[0] + yo[i] * e1[1] + zo[i] * e1[2] + sx;
            double oy = xo[i] * e2[0] + yo[i] * e2[1] + zo[i] * e2[2] + sy;
            double oz = xo[i] * e3[0] + yo[i] * e3[1] + zo[i] * e3[2] + sz;
            xo[i] = ox; yo[i] = oy; zo[i] = oz;
            double wx = vx[i] * e1[0] + vy[i] * e1[1] + vz[i] * e1[2];
            double wy = vx[i] * e2[0] + vy[i] * e2[1] + vz[i] * e2[2];
            double wz = vx[i] * e3[0] + vy[i] * e3[1] + vz[i] * e3[2];
            vx[i] = wx; vy[i] = wy; vz[i] = wz;
        }
    }

    // Eigenvectors of the inertia tensor [[A,-H,-G],[-H,B,-F],[-G,-F,C]] as the
    // reference computes them: closed forms when two products vanish, the
    // trigonometric cubic solution otherwise; e3 = e1 x e2.
    static void principal_axes(double A, double B, double C, double F, double G, double H,
                               double *e1, double *e2, double *e3) {
        double n1[3] = {0, 0, 0}, n2[3] = {0, 0, 0}, n3[3] = {0, 0, 0};
        double *p1 = nullptr, *p2 = nullptr;
        const double tiny = 1.0e-6;
        auto order3 = [&](double l1, double l2, double l3) {
            // sort eigenpairs (l1,n1) (l2,n2) (l3,n3) ascending; p1,p2 = two smallest
            double lo, mid;
            double *q1, *q2;
            if (l1 < l2) { lo = l1; mid = l2; q1 = n1; q2 = n2; }
            else { lo = l2; mid = l1; q1 = n2; q2 = n1; }
            if (l3 < lo) { q2 = q1; q1 = n3; }
            else if (l3 < mid) { q2 = n3; }
            p1 = q1; p2 = q2;
            if (p1[0] < 0) { p1[0] = -p1[0]; p1[1] = -p1[1]; p1[2] = -p1[2]; }
            if (p2[1] < 0) { p2[0] = -p2[0]; p2[1] = -p2[1]; p2[2] = -p2[2]; }
        };
        auto norm = [](double *v) {
            double s = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            v[0] /= s; v[1] /= s; v[2] /= s;
        };
        if (std::fabs(F) < tiny && std::fabs(G) < tiny) {
            double s = std::sqrt(B * B - 2.0 * A * B + A * A + 4.0 * H * H);
            double l1 = 0.5 * (B + A + s);
            n1[0] = 0.5 * (B - A - s) / H; n1[1] = 1; n1[2] = 0; norm(n1);
            double l2 = 0.5 * (B + A - s);
            n2[0] = 0.5 * (B - A + s) / H; n2[1] = 1; n2[2] = 0; norm(n2);
            n3[2] = 1;
            order3(l1, l2, C);
        } else if (std::fabs(F) < tiny && std::fabs(H) < tiny) {
            double s = std::sqrt(C * C - 2.0 * A * C + A * A + 4.0 * G * G);
            double l1 = 0.5 * (C + A + s);
            n1[0] = 0.5 * (C - A - s) / G; n1[1] = 0; n1[2] = 1; norm(n1);
            double l2 = 0.5 * (C + A - s);
            n2[0] = 0.5 * (C - A + s) / G; n2[1] = 0; n2[2] = 1; norm(n2);
            n3[1] = 1;
            order3(l1, l2, B);
        } else if (std::fabs(G) < tiny && std::fabs(H) < tiny) {
            double s = std::sqrt(C * C - 2.0 * B * C + B * B + 4.0 * F * F);
            double l1 = 0.5 * (C + B + s);
            n1[1] = 0.5 * (C - B - s) / F; n1[0] = 0; n1[2] = 1; norm(n1);
            double l2 = 0.5 * (C + B - s);
            n2[1] = 0.5 * (C - B + s) / F; n2[0] = 0; n2[2] = 1; norm(n2);
            n3[0] = 1;
            order3(l1, l2, A);
        } else {
            // Roots of det(I - l) = l^3 + c2 l^2 + c1 l + c0 by Viete's trigonometric method.
            double c2 = -A - B - C;
            double c1 = A * B + A * C + B * C - H * H - G * G - F * F;
            double c0 = A * F * F + B * G * G + C * H * H + 2.0 * H * G * F - A * B * C;
            double m = -c2 / 3.0;
            double q = (9.0 * c1 * c2 - 27 * c0 - 2.0 * c2 * c2 * c2) / 54.0;
            double p = m * m - c1 / 3.0;
            double rad = 2.0 * std::sqrt(p);
            double phi = std::acos(q * std::pow(p, -1.5)) / 3.0;
            double third = kPi / 3.0;
            double l1 = m + rad * std::cos(phi);
            double l2 = m - rad * std::cos(phi - third);
            double l3 = m - rad * std::cos(phi + third);
            double lo, mid;
            if (l1 < l2) { lo = l1; mid = l2; } else { lo = l2; mid = l1; }
            if (l3 < lo) { mid = lo; lo = l3; }
            else if (l3 < mid) { mid = l3; }
            double u = (A - lo) * F + G * H, v = (B - lo) * G + F * H, w = (C - lo) * H + G * F;
            double r1 = u / v, r2 = u / w;
            double s1 = std::sqrt(1.0 + r1 * r1 + r2 * r2);
            n1[0] = 1.0 / s1; n1[1] = r1 / s1; n1[2] = r2 / s1;
            u = (A - mid) * F + G * H; v = (B - mid) * G + F * H; w = (C - mid) * H + G * F;
            r1 = v / u; r2 = v / w;
            s1 = std::sqrt(1.0 + r1 * r1 + r2 * r2);
            n2[0] = r1 / s1; n2[1] = 1.0 / s1; n2[2] = r2 / s1;
            p1 = n1; p2 = n2;
        }
        for (int c = 0; c < 3; ++c) { e1[c] = p1[c]; e2[c] = p2[c]; }
        cross(e1, e2, e3);
    }

    // add_crowder for the three single-site ion species (fast.cpp:5055-5403):
    // rejection-sample a uniform position that overlaps nothing placed so far.
    void add_ion(int species) {
        static const int kCls[3] = {3, 4, 5}, kType[3] = {7, 8, 9};
        int k = new_particle(kCls[species], kType[species]);
        int tk = maxi[k];
        for (;;) {
            double px = uniform() * box;
            double py = uniform() * box;
            double pz = uniform() * box;
            // Orientation frame of the (single-site) ion: drawn, never used.
            for (int q = 0; q < 12; ++q) rng.next();
            x[k] = px; y[k] = py; z[k] = pz;
            bool clash = false;
            for (int j = 1; j < k && !clash; ++j) {
                double d[3] = {px - x[j], py - y[j], pz - z[j]};
                for (int c = 0; c < 3; ++c) {
                    if (d[c] > half[c]) d[c] -= box;
                    if (d[c] < -half[c]) d[c] += box;
                }
                double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (r2 < d2_overlap(tk, maxi[j])) clash = true;
            }
            if (!clash) break;
        }
        draw_velocity(k);
    }

    void place_ions() {
        n_ion[0] = (int)(args.c_mg * 6.022 * 0.0001 * volume);
        n_ion[2] = (int)(args.c_k * 6.022 * 0.0001 * volume);
        n_ion[1] = 2 * n_ion[0] + n_ion[2] - (n_units - 1) / 2;
        if (n_ion[1] < 0) throw IoError("negative chloride count: salt too low to neutralise");
        for (int s = 0; s < 3; ++s)
            for (int q = 0; q < n_ion[s]; ++q) add_ion(s);
    }
};

// Backbone bond/angle lists in the order polymer_constraints visits them
// (fast.cpp:2360-2519): base bond + two B-S-P angles per nucleoside, then
// S(n)-P(n+1) bonds with P-S-P angles, then P-S bonds with S-P-S angles.
// Base bond and B-S-P angles per base: r0, theta0 towards P(n+1), theta0 towards P(n).
struct BaseParam { char k; double r0, a1, a2; };
const BaseParam kBase[4] = {{'A', 4.8515, 1.9259, 1.7029},
                            {'C', 4.2738, 1.9655, 1.5803},
                            {'G', 4.9659, 1.9150, 1.7690},
                            {'U', 4.2733, 1.9663, 1.5735}};
const BaseParam *base_param(char k) {
    for (const BaseParam &q : kBase)
        if (q.k == k) return &q;
    return nullptr;
}
// Sugar-phosphate backbone: S(n)-P(n+1) and P(n)-S(n) bond lengths, the
// P-S-P angle at S and the S-P-S angle at P.
constexpr double kBondSP = 3.8157, kBondPS = 4.6010, kAngleS = 1.4440, kAngleP = 1.5256;

void backbone_terms(const Builder &b, std::vector<int> &bij, std::vector<double> &bkr,
                    std::vector<int> &aijk, std::vector<double> &akt) {
    int n = b.n_units;
    for (int e = 1; e <= n; e += 2) {
        int i = b.key0[e], j = b.key1[e];
        const BaseParam *p = base_param(b.kind[e]);
        if (!p) continue;
        bij.insert(bij.end(), {i - 1, j - 1});
        bkr.insert(bkr.end(), {10.0, p->r0});
        if (e < n) {
            int k = b.key0[e + 1];
            aijk.insert(aijk.end(), {j - 1, i - 1, k - 1});
            akt.insert(akt.end(), {5.0, p->a1});
        }
        if (e > 1) {
            int k = b.key0[e - 1];
            aijk.insert(aijk.end(), {j - 1, i - 1, k - 1});
            akt.insert(akt.end(), {5.0, p->a2});
        }
    }
    for (int e = 2; e <= n; e += 2) {
        int i = b.key0[e], j = b.key0[e - 1];
        bij.insert(bij.end(), {i - 1, j - 1});
        bkr.insert(bkr.end(), {64.0, kBondSP});
        if (e > 2) {
            int k = b.key0[e - 2];
            aijk.insert(aijk.end(), {i - 1, j - 1, k - 1});
            akt.insert(akt.end(), {20.0, kAngleS});
        }
    }
    for (int e = 3; e <= n; e += 2) {
        int i = b.key0[e], j = b.key0[e - 1], k = b.key0[e - 2];
        bij.insert(bij.end(), {i - 1, j - 1});
        bkr.insert(bkr.end(), {23.0, kBondPS});
        aijk.insert(aijk.end(), {i - 1, j - 1, k - 1});
        akt.insert(akt.end(), {20.0, kAngleP});
    }
}

template <class T>
T *dup(const std::vector<T> &v) {
    T *p = (T *)std::malloc(sizeof(T) * (v.empty() ? 1 : v.size()));
    if (!p) throw std::bad_alloc();
    if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

tis_system *export_system(const Builder &b) {
    tis_system *s = (tis_system *)std::calloc(1, sizeof(tis_system));
    if (!s) throw std::bad_alloc();
    int N = (int)b.x.size() - 1;
    s->n_particles = N;
    s->n_beads = b.n_beads;
    s->n_units = b.n_units;
    s->n_mg = b.n_ion[0]; s->n_cl = b.n_ion[1]; s->n_k = b.n_ion[2];
    s->box[0] = s->box[1] = s->box[2] = b.box;
    s->T = b.args.T;
    s->lB = 332.0637090 * (0.03273600947 * b.args.T - 0.00669625750);
    s->h = 0.05;
    s->recentre = b.kh - 1;
    for (int t = 0; t < TIS_NTYPES; ++t) {
        s->type_mass[t] = b.mass[t + 1]; s->type_R[t] = b.rlj[t + 1];
        s->type_eps[t] = b.elj[t + 1]; s->type_q[t] = b.qc[t + 1];
    }
    std::vector<int> type(N), cls(N);
    std::vector<double> pos(3 * N), old(3 * N), vel(3 * N);
    for (int i = 1; i <= N; ++i) {
        type[i - 1] = b.maxi[i] - 1;
        cls[i - 1] = b.part[i];
        pos[3 * i - 3] = b.x[i]; pos[3 * i - 2] = b.y[i]; pos[3 * i - 1] = b.z[i];
        old[3 * i - 3] = b.xo[i]; old[3 * i - 2] = b.yo[i]; old[3 * i - 1] = b.zo[i];
        vel[3 * i - 3] = b.vx[i]; vel[3 * i - 2] = b.vy[i]; vel[3 * i - 1] = b.vz[i];
    }
    s->type = dup(type); s->cls = dup(cls);
    s->pos = dup(pos); s->pos_old = dup(old); s->vel = dup(vel);
    std::vector<int> u0(b.n_units), u1(b.n_units);
    std::vector<char> uk(b.n_units);
    for (int u = 1; u <= b.n_units; ++u) {
        u0[u - 1] = b.key0[u] - 1;
        u1[u - 1] = b.kind[u] == 'P' ? -1 : b.key1[u] - 1;
        uk[u - 1] = b.kind[u];
    }
    s->unit_bead0 = dup(u0); s->unit_bead1 = dup(u1); s->unit_kind = dup(uk);
    std::vector<int> bij, aijk;
    std::vector<double> bkr, akt;
    backbone_terms(b, bij, bkr, aijk, akt);
    s->n_bonds = (int)bkr.size() / 2;
    s->n_angles = (int)akt.size() / 2;
    s->bond_ij = dup(bij); s->bond_kr = dup(bkr);
    s->angle_ijk = dup(aijk); s->angle_kt = dup(akt);
    // tertiary stacks
    int n3 = (int)b.st3.size() - 1, n2 = (int)b.st2.size() - 1;
    s->n_st3 = n3; s->n_st2 = n2;
    std::vector<int> sb(6 * n3), sf(2 * n3), sd(2 * n3);
    std::vector<double> sg(6 * n3), se(n3);
    for (int k = 1; k <= n3; ++k) {
        const Builder::St3 &t = b.st3[k];
        for (int c = 0; c < 6; ++c) { sb[6 * k - 6 + c] = t.b[c] - 1; sg[6 * k - 6 + c] = t.g[c]; }
        se[k - 1] = t.E;
        for (int c = 0; c < 2; ++c) {
            int f = t.face[c];
            sf[2 * k - 2 + c] = f;
            // A formed tertiary stack switches off the secondary stack on a face
            // shared by exactly two stacks: index s3_N + ceil(f/2) - 1 (fast.cpp:1960-1981).
            int m = (f + 1) / 2 - 2;   // 0-based secondary index
            sd[2 * k - 2 + c] = (b.st_excess[f] == 1 && m >= 0 && m < n2) ? m : -1;
        }
    }
    s->st3_beads = dup(sb); s->st3_geo = dup(sg); s->st3_E = dup(se);
    s->st3_face = dup(sf); s->st3_disable = dup(sd);
    std::vector<int> s2b(7 * n2);
    std::vector<double> s2r(n2), s2e(n2);
    for (int k = 1; k <= n2; ++k) {
        int i = b.st2[k].unit;
        int beads[7] = {b.key0[i - 1], b.key0[i], b.key1[i], b.key0[i + 1],
                        b.key0[i + 2], b.key1[i + 2], b.key0[i + 3]};
        for (int c = 0; c < 7; ++c) s2b[7 * k - 7 + c] = beads[c] - 1;
        s2r[k - 1] = b.st2[k].r0;
        s2e[k - 1] = b.st2[k].E;
    }
    s->st2_beads = dup(s2b); s->st2_r0 = dup(s2r); s->st2_E = dup(s2e);
    std::vector<int> sx(b.st_excess.begin() + 1, b.st_excess.end());
    s->st_excess = dup(sx);
    // hydrogen bonds
    int nh = (int)b.hb.size() - 1, ns = (int)b.valence.size() - 1;
    s->n_hb = nh; s->n_hb_sites = ns;
    std::vector<int> hbb(6 * nh), hd(nh), hso(nh + 1), hsi, sho(ns + 1), shi, val(ns);
    std::vector<double> hg(6 * nh), he(nh);
    for (int k = 1; k <= nh; ++k) {
        const Builder::Hb &h = b.hb[k];
        for (int c = 0; c < 6; ++c) { hbb[6 * k - 6 + c] = h.b[c] - 1; hg[6 * k - 6 + c] = h.g[c]; }
        he[k - 1] = h.E;
        hd[k - 1] = h.dode;
        hso[k - 1] = (int)hsi.size();
        for (int st : h.sites) hsi.push_back(st - 1);
    }
    hso[nh] = (int)hsi.size();
    for (int q = 1; q <= ns; ++q) {
        sho[q - 1] = (int)shi.size();
        for (int k : b.site_hbs[q]) shi.push_back(k - 1);
        val[q - 1] = b.valence[q];
    }
    sho[ns] = (int)shi.size();
    s->n_hb_incid = (int)hsi.size();
    s->hb_beads = dup(hbb); s->hb_geo = dup(hg); s->hb_E = dup(he); s->hb_dode = dup(hd);
    s->hb_site_off = dup(hso); s->hb_site_idx = dup(hsi);
    s->site_hb_off = dup(sho); s->site_hb_idx = dup(shi); s->site_valence = dup(val);
    return s;
}

// ------------------------------------------------------------ synthetic chains
//
// The reference builds only 16SCD, from its PDB and hand-written HB/stack
// lists with hard-coded residue offsets (fast.cpp:3137-4492).  A synthetic
// chain of any length gets the same tables from a self-avoiding walk instead:
// units and beads in the reference's order (S1 B1 P2 S2 B2 ...), every bond
// and bond angle at the equilibrium value of its backbone term, secondary
// stacks on a random subset of adjacent nucleotide pairs (same parameter
// table as initialize_unprocessed_stacks), and HB candidates between bases
// the walk brings into contact, their native geometry measured on the built
// chain the way initialize_unprocessed_bonds measures it on the PDB.

Vec3 add(const Vec3 &a, const Vec3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec3 sub(const Vec3 &a, const Vec3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec3 scale(const Vec3 &a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(const Vec3 &a, const Vec3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3 crossv(const Vec3 &a, const Vec3 &b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
Vec3 unit(const Vec3 &a) { return scale(a, 1.0 / std::sqrt(dot(a, a))); }

// Closest approach the walk allows between beads that are neither bonded nor
// 1-3 partners (the RNA-RNA WCA repels below D_LJ = 3.2 A).
constexpr double kClash = 3.5;

class ChainGrower {
  public:
    ChainGrower(Builder &b, double radius, double clash)
        : b_(b), radius_(radius), clash_(clash), clash2_(clash * clash) {}

    // Sequence, units and beads, then positions.  Step n places S(n) (n > 1),
    // P(n+1) (n < N) and B(n); a step that finds no free position undoes
    // itself and the walk backs up, further each time it fails again before
    // reaching new ground.
    void grow(int n_nt) {
        static const char kSeq[4] = {'A', 'C', 'G', 'U'};
        std::vector<char> seq(n_nt + 1);
        for (int n = 1; n <= n_nt; ++n) seq[n] = kSeq[std::min(3, (int)(b_.uniform() * 4.0))];
        b_.kind.assign(1, 0);
        for (int n = 1; n <= n_nt; ++n) {
            if (n > 1) b_.kind.push_back('P');
            b_.kind.push_back(seq[n]);
        }
        b_.n_units = (int)b_.kind.size() - 1;
        b_.key0.assign(b_.n_units + 1, 0);
        b_.key1.assign(b_.n_units + 1, 0);
        for (int u = 1; u <= b_.n_units; ++u) {
            if (b_.kind[u] == 'P') {
                b_.key0[u] = b_.new_particle(0, 1);
                b_.key1[u] = -1;
            } else {
                char k = b_.kind[u];
                b_.key0[u] = b_.new_particle(1, 2);
                b_.key1[u] = b_.new_particle(2, k == 'A' ? 3 : k == 'G' ? 4 : k == 'C' ? 5 : 6);
            }
        }
        b_.n_beads = (int)b_.x.size() - 1;
        n_nt_ = n_nt;

        put(S(1), {0, 0, 0});
        std::vector<size_t> mark(n_nt + 2, 0);
        budget_ = 2000LL * n_nt;
        int n = 1, best = 0, depth = 1;
        while (n <= n_nt) {
            mark[n] = placed_.size();
            if (step(n)) {
                if (n > best) { best = n; depth = 1; }
                ++n;
                continue;
            }
            undo_to(mark[n]);
            if (budget_ < 0)
                throw IoError("synthetic chain: the walk found no free position after nucleotide " +
                              std::to_string(best) + " (radius too small)");
            int back = std::min(depth, n - 1);
            n -= back;
            undo_to(mark[n]);
            depth = std::min(2 * depth, 64);
        }
    }

  private:
    Builder &b_;
    double radius_, clash_, clash2_;
    int n_nt_ = 0;
    long long budget_ = 0;        // trial positions left before the walk gives up
    std::unordered_map<long long, std::vector<int>> grid_;
    std::vector<int> placed_;

    int S(int n) const { return b_.key0[2 * n - 1]; }
    int B(int n) const { return b_.key1[2 * n - 1]; }
    int Pn(int n) const { return b_.key0[2 * n - 2]; }   // n >= 2

    // The base's position follows from S(n) and the two phosphates, so a
    // blocked base is retried with new dihedrals for S(n) and P(n+1).
    bool step(int n) {
        for (int t = 0; t < 32; ++t) {
            size_t m = placed_.size();
            if (n > 1 && !place_backbone(S(n), n > 2 ? Pn(n - 1) : -1, S(n - 1), Pn(n), kBondPS, kAngleP,
                                         {Pn(n), S(n - 1)}))
                return false;
            if (n < n_nt_ && !place_backbone(Pn(n + 1), n > 1 ? S(n - 1) : -1, n > 1 ? Pn(n) : -1, S(n),
                                             kBondSP, kAngleS, {S(n), n > 1 ? Pn(n) : -1})) {
                undo_to(m);
                continue;
            }
            if (place_base(n, S(n), B(n), n > 1 ? Pn(n) : -1, n < n_nt_ ? Pn(n + 1) : -1)) return true;
            undo_to(m);
        }
        return false;
    }

    long long cell_key(int cx, int cy, int cz) const {
        const long long o = 1 << 20;
        return ((cx + o) << 42) | ((cy + o) << 21) | (cz + o);
    }
    int cell_of(double v) const { return (int)std::floor(v / clash_); }
    long long key_of(int k) const { return cell_key(cell_of(b_.x[k]), cell_of(b_.y[k]), cell_of(b_.z[k])); }
    void put(int k, const Vec3 &p) {
        b_.x[k] = p.x; b_.y[k] = p.y; b_.z[k] = p.z;
        grid_[key_of(k)].push_back(k);
        placed_.push_back(k);
    }
    void undo_to(size_t mark) {
        while (placed_.size() > mark) {
            int k = placed_.back();
            placed_.pop_back();
            std::vector<int> &c = grid_[key_of(k)];
            c.erase(std::find(c.begin(), c.end(), k));
        }
    }
    // True when p is closer than the clash distance to a placed bead other
    // than the listed bonded / 1-3 partners.
    bool clashes(const Vec3 &p, std::initializer_list<int> skip) {
        if (--budget_ < 0) return true;
        int cx = cell_of(p.x), cy = cell_of(p.y), cz = cell_of(p.z);
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    auto it = grid_.find(cell_key(cx + dx, cy + dy, cz + dz));
                    if (it == grid_.end()) continue;
                    for (int j : it->second) {
                        bool skipped = false;
                        for (int s : skip) skipped |= s == j;
                        if (skipped) continue;
                        Vec3 d = sub(p, b_.P(j));
                        if (dot(d, d) < clash2_) return true;
                    }
                }
        return false;
    }
    Vec3 random_unit() {
        double z = 2.0 * b_.uniform() - 1.0, ph = 2.0 * kPi * b_.uniform();
        double s = std::sqrt(std::max(0.0, 1.0 - z * z));
        return {s * std::cos(ph), s * std::sin(ph), z};
    }
    // Orthonormal pair perpendicular to the unit vector u.
    void frame(const Vec3 &u, Vec3 &e1, Vec3 &e2) {
        Vec3 t = std::fabs(u.x) < 0.9 ? Vec3{1, 0, 0} : Vec3{0, 1, 0};
        e1 = unit(crossv(u, t));
        e2 = crossv(u, e1);
    }
    bool inside(const Vec3 &p, double r) const { return dot(p, p) <= r * r; }

    // Bead k at distance r from c, with angle theta at c towards bd and a
    // random dihedral about (bd, c); a is the bead before bd, or -1 at the
    // chain start (then the dihedral reference is a random direction).
    // 64 random dihedrals; the last 16 may leave the sphere by up to 25%.
    bool place_backbone(int k, int a, int bd, int c, double r, double theta,
                        std::initializer_list<int> skip) {
        Vec3 C = b_.P(c);
        for (int t = 0; t < 64; ++t) {
            Vec3 p;
            if (bd < 0) {
                p = add(C, scale(random_unit(), r));
            } else {
                Vec3 Bv = b_.P(bd);
                Vec3 bc = unit(sub(C, Bv));
                Vec3 ref = a >= 0 ? sub(Bv, b_.P(a)) : random_unit();
                Vec3 nrm = crossv(ref, bc);
                if (dot(nrm, nrm) < 1e-12) continue;
                nrm = unit(nrm);
                Vec3 m = crossv(nrm, bc);
                double ph = 2.0 * kPi * b_.uniform();
                Vec3 d = add(scale(bc, -r * std::cos(theta)),
                             add(scale(m, r * std::sin(theta) * std::cos(ph)),
                                 scale(nrm, r * std::sin(theta) * std::sin(ph))));
                p = add(C, d);
            }
            if (!inside(p, t < 48 ? radius_ : 1.25 * radius_) || clashes(p, skip)) continue;
            put(k, p);
            return true;
        }
        return false;
    }

    // Base bead of nucleotide n: bond to S at the base's r0 with both B-S-P
    // angles at their equilibrium values where both phosphates exist (two
    // mirror solutions), else on the cone of the one angle that exists.
    bool place_base(int n, int s, int bk, int p_prev, int p_next) {
        const BaseParam *bp = base_param(b_.kind[2 * n - 1]);
        Vec3 Sv = b_.P(s);
        if (p_prev >= 0 && p_next >= 0) {
            Vec3 u1 = unit(sub(b_.P(p_prev), Sv)), u2 = unit(sub(b_.P(p_next), Sv));
            double c = dot(u1, u2), c1 = std::cos(bp->a2), c2 = std::cos(bp->a1);
            double den = 1.0 - c * c;
            if (!(den > 1e-9)) return false;
            double al = (c1 - c * c2) / den, be = (c2 - c * c1) / den;
            Vec3 inplane = add(scale(u1, al), scale(u2, be));
            Vec3 nrm = crossv(u1, u2);
            double g2 = (1.0 - dot(inplane, inplane)) / dot(nrm, nrm);
            if (g2 < 0) return false;
            double g = std::sqrt(g2);
            bool flip = b_.uniform() < 0.5;
            for (int q = 0; q < 2; ++q) {
                Vec3 dir = add(inplane, scale(nrm, (flip != (q == 1)) ? -g : g));
                Vec3 p = add(Sv, scale(dir, bp->r0));
                if (inside(p, 1.25 * radius_) && !clashes(p, {s, p_prev, p_next})) { put(bk, p); return true; }
            }
            return false;
        }
        // Chain ends: on the cone about the one phosphate direction.
        int axis_bead = p_next >= 0 ? p_next : p_prev;
        double ang = p_next >= 0 ? bp->a1 : bp->a2;
        Vec3 u = unit(sub(b_.P(axis_bead), Sv)), e1, e2;
        frame(u, e1, e2);
        for (int t = 0; t < 64; ++t) {
            double ps = 2.0 * kPi * b_.uniform();
            Vec3 dir = add(scale(u, std::cos(ang)),
                           add(scale(e1, std::sin(ang) * std::cos(ps)), scale(e2, std::sin(ang) * std::sin(ps))));
            Vec3 p = add(Sv, scale(dir, bp->r0));
            if (!inside(p, 1.25 * radius_) || clashes(p, {s, p_prev, p_next})) continue;
            put(bk, p);
            return true;
        }
        return false;
    }
};

// Secondary stacks on a random subset of adjacent pairs of units (i, i+2),
// i = 3, 5, ... as initialize_unprocessed_stacks enumerates them
// (fast.cpp:4399-4429); ST_EXCESS as there, with no tertiary stacks.
void synthetic_stacks(Builder &b, double frac) {
    b.st_excess.assign(b.n_units + 1, 0);
    for (int i = 3; i < b.n_units - 2; i += 2) {
        if (!(b.uniform() < frac)) continue;
        const SecStackParam *p = nullptr;
        for (const SecStackParam &q : kSecStack)
            if (q.a == b.kind[i] && q.b == b.kind[i + 2]) p = &q;
        Builder::St2 s{};
        s.unit = i;
        s.r0 = p->r0;
        s.E = -p->dG + b.args.ss_D / p->c + p->s * (b.args.T - p->Tm);
        b.st2.push_back(s);
        b.st_excess[i] += 1;
        b.st_excess[i + 1] += 1;
    }
    for (int f = 1; f <= b.n_units; ++f) b.st_excess[f] -= 1;
}

// HB candidates between nucleotides i < j (j >= i + 4, both with a 3'
// phosphate) whose base beads lie r_lo..r_hi apart: pairs drawn in random
// order, each nucleotide in at most two candidates, up to n_hb.  A candidate
// joins 2 atom-level bonds (3 for G-C) between distinct base sites of the two
// nucleotides, E = -(bonds) hs_D kHbUnit as summed in
// initialize_unprocessed_bonds (fast.cpp:3881-3890); tertiary class when the
// nucleotides are more than 50 apart in sequence.
void synthetic_hbonds(Builder &b, int n_nt, int n_hb, double r_lo, double r_hi) {
    auto S = [&](int n) { return b.key0[2 * n - 1]; };
    auto Bb = [&](int n) { return b.key1[2 * n - 1]; };
    auto P3 = [&](int n) { return b.key0[2 * n]; };
    std::unordered_map<long long, std::vector<int>> cells;
    auto key = [](int cx, int cy, int cz) {
        const long long o = 1 << 20;
        return ((cx + o) << 42) | ((cy + o) << 21) | (cz + o);
    };
    auto cidx = [&](double v) { return (int)std::floor(v / r_hi); };
    for (int n = 1; n < n_nt; ++n) {
        Vec3 p = b.P(Bb(n));
        cells[key(cidx(p.x), cidx(p.y), cidx(p.z))].push_back(n);
    }
    std::vector<std::pair<int, int>> pairs;
    for (int i = 1; i < n_nt; ++i) {
        Vec3 p = b.P(Bb(i));
        int cx = cidx(p.x), cy = cidx(p.y), cz = cidx(p.z);
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    auto it = cells.find(key(cx + dx, cy + dy, cz + dz));
                    if (it == cells.end()) continue;
                    for (int j : it->second) {
                        if (j < i + 4) continue;
                        double r = dist(p, b.P(Bb(j)));
                        if (r >= r_lo && r <= r_hi) pairs.emplace_back(i, j);
                    }
                }
    }
    std::sort(pairs.begin(), pairs.end());
    for (int q = (int)pairs.size() - 1; q > 0; --q) {
        int r = std::min(q, (int)(b.uniform() * (q + 1)));
        std::swap(pairs[q], pairs[r]);
    }
    std::vector<int> uses(n_nt + 1, 0);
    std::vector<std::pair<int, int>> take;
    for (const auto &pr : pairs) {
        if ((int)take.size() >= n_hb) break;
        if (uses[pr.first] >= 2 || uses[pr.second] >= 2) continue;
        ++uses[pr.first];
        ++uses[pr.second];
        take.push_back(pr);
    }
    std::sort(take.begin(), take.end());
    auto pick_slots = [&](int unit_idx, int m, int *out) {
        // base sites are slots 2..ns of the nucleotide (slot 0 O2', 1 O4')
        std::vector<int> slots;
        for (int q = 2; q < (int)b.unit_sites[unit_idx].size(); ++q) slots.push_back(q);
        for (int q = 0; q < m; ++q) {
            int r = q + std::min((int)slots.size() - 1 - q, (int)(b.uniform() * (slots.size() - q)));
            std::swap(slots[q], slots[r]);
            out[q] = slots[q];
        }
    };
    for (const auto &pr : take) {
        int i = pr.first, j = pr.second;
        char ki = b.kind[2 * i - 1], kj = b.kind[2 * j - 1];
        int m = ((ki == 'G' && kj == 'C') || (ki == 'C' && kj == 'G')) ? 3 : 2;
        Builder::Hb h{};
        int bb[6] = {S(i), Bb(i), P3(i), S(j), Bb(j), P3(j)};
        std::memcpy(h.b, bb, sizeof bb);
        h.g[0] = dist(b.P(bb[1]), b.P(bb[4]));
        h.g[1] = angle(b.P(bb[0]), b.P(bb[1]), b.P(bb[4]));
        h.g[2] = angle(b.P(bb[3]), b.P(bb[4]), b.P(bb[1]));
        h.g[3] = dihedral(b.P(bb[0]), b.P(bb[1]), b.P(bb[4]), b.P(bb[3]));
        h.g[4] = dihedral(b.P(bb[4]), b.P(bb[1]), b.P(bb[0]), b.P(bb[2]));
        h.g[5] = dihedral(b.P(bb[1]), b.P(bb[4]), b.P(bb[3]), b.P(bb[5]));
        h.E = -m * b.args.hs_D * kHbUnit;
        h.dode = j - i > 50 ? 1 : 0;
        int si[3], sj[3];
        pick_slots(2 * i - 1, m, si);
        pick_slots(2 * j - 1, m, sj);
        int k = (int)b.hb.size();
        for (int q = 0; q < m; ++q) {
            int s1 = b.unit_sites[2 * i - 1][si[q]], s2 = b.unit_sites[2 * j - 1][sj[q]];
            h.sites.push_back(s1);
            h.sites.push_back(s2);
            b.site_hbs[s1].push_back(k);
            b.site_hbs[s2].push_back(k);
        }
        b.hb.push_back(h);
    }
}

}  // namespace

tis_system *build_chain(const tis_chain_files &f, const tis_build_args &args, long long *draws) {
    Builder b;
    b.args = args;
    b.box = args.box > 0 ? args.box : 350.0;
    b.kh = f.recentre_bead > 0 ? f.recentre_bead : 500;
    b.rng.reseed((unsigned)(args.T * 1000 * args.seed));
    b.read_maxi_key(f.maxikey);
    b.read_pdb(f.pdb);
    b.init_hb_sites();
    b.read_hb_candidates(f.bonds);
    std::vector<int> faces;
    if (f.double_face_res) faces.assign(f.double_face_res, f.double_face_res + f.n_double_face);
    b.read_stacks(f.stacks, faces);
    if (b.n_beads < b.kh)
        throw IoError("chain of " + std::to_string(b.n_beads) + " beads is shorter than the recentring bead " +
                      std::to_string(b.kh));
    b.set_box(b.kh);
    b.place_ions();
    if (draws) *draws = b.rng.draws();
    return export_system(b);
}

tis_system *build_16scd(const std::string &dir, const tis_build_args &args, long long *draws) {
    std::string d = dir.empty() ? std::string(".") : dir;
    std::string pdb = d + "/16SCD.pdb", bonds = d + "/16SCD_16SCD_unprocessed_bonds.dat",
                stacks = d + "/16SCD_16SCD_unprocessed_stacks.dat", maxi = d + "/MAXIKEY_Mg_Cl_K.dat";
    tis_chain_files f{pdb.c_str(), bonds.c_str(), stacks.c_str(), maxi.c_str(), k16scdDoubleFace, 15, 500};
    return build_chain(f, args, draws);
}

tis_system *build_synthetic(const std::string &dir, const tis_synthetic_args &a) {
    Builder b;
    b.args = a.build;
    b.box = a.build.box;
    b.rng.reseed((unsigned)(a.build.T * 1000 * a.build.seed));
    b.read_maxi_key((dir.empty() ? std::string(".") : dir) + "/MAXIKEY_Mg_Cl_K.dat");
    ChainGrower grower(b, a.radius, kClash);
    grower.grow(a.n_nt);
    b.init_hb_sites();
    synthetic_stacks(b, a.stack_frac);
    synthetic_hbonds(b, a.n_nt, a.n_hb, 5.0, 7.0);
    // velocities unit by unit, S before B, as read_PDB draws them (fast.cpp:3625-3627)
    for (int u = 1; u <= b.n_units; ++u) {
        b.draw_velocity(b.key0[u]);
        if (b.kind[u] != 'P') b.draw_velocity(b.key1[u]);
    }
    b.kh = (b.n_beads + 1) / 2;
    b.set_box(b.kh);
    b.place_ions();
    return export_system(b);
}

}  // namespace tis_host
