// Device-side arithmetic shared by the engine kernels: small vectors, the
// minimum-image rule, the bead geometry and the harmonic force primitives of
// the TIS model, templated on the force precision R (float or double).
//
// Positions are always FP64; a difference of two positions is formed in FP64
// and only then rounded to R, so FP32 mode loses nothing to the absolute size
// of the coordinates (0..600 A).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TIS_DEV __device__ __forceinline__
#define TIS_HD __host__ __device__ __forceinline__

namespace tis {

constexpr double kPi = 3.14159265358979323846264338328;
constexpr double kSigLJ = 1.5852;          // WCA shift length, fast.cpp:801-821
constexpr double kCoulCut2 = 2500.0;       // D2_CT_CUTOFF, fast.cpp:348
constexpr double kRandMax = 2147483647.0;  // glibc RAND_MAX

template <class R>
struct V3 {
    R x, y, z;
};
template <class R>
TIS_DEV V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R>
TIS_DEV V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R>
TIS_DEV V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R>
TIS_DEV R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R>
TIS_DEV V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

TIS_DEV double rsqrt_(double v) { return 1.0 / sqrt(v); }
TIS_DEV float rsqrt_(float v) { return 1.0f / sqrtf(v); }

// Difference b - a of two FP64 positions, rounded to R (no periodic image:
// bonded geometry never spans the box, as in the reference).
template <class R>
TIS_DEV V3<R> diff(const double4 &a, const double4 &b) {
    return {(R)(b.x - a.x), (R)(b.y - a.y), (R)(b.z - a.z)};
}

// Squared minimum-image distance with the reference's exact FP64 operation
// order (half_shift then dx*dx + dy*dy + dz*dz, fast.cpp:646-653, 1620-1644);
// the _rn intrinsics forbid contraction into FMAs so that list membership is
// decided by the same rounded value the reference compares.
TIS_DEV double min_image_r2(const double4 &pi, const double4 &pj, const double *L, const double *hL,
                            double3 *d_out = nullptr) {
    double d[3] = {__dsub_rn(pi.x, pj.x), __dsub_rn(pi.y, pj.y), __dsub_rn(pi.z, pj.z)};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (d[c] > hL[c]) d[c] = __dsub_rn(d[c], L[c]);
        if (d[c] < -hL[c]) d[c] = __dadd_rn(d[c], L[c]);
    }
    if (d_out) *d_out = make_double3(d[0], d[1], d[2]);
    return __dadd_rn(__dadd_rn(__dmul_rn(d[0], d[0]), __dmul_rn(d[1], d[1])), __dmul_rn(d[2], d[2]));
}

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of
// (hi, lo) = M * c on two lanes, xor with the other lanes and the key,
// key bumped by the Weyl constants between rounds.
struct Philox4 {
    uint32_t v[4];
};
TIS_HD Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * c0;
        uint64_t p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    Philox4 o;
    o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
    return o;
}
// Uniform in [0, 1) with 32-bit resolution (HB draws: only ceil(u*n) and u < p are used).
TIS_HD double u01_closed_open(uint32_t r) { return (double)r * 2.3283064365386963e-10; }
// Uniform in (0, 1] for the logarithm of Box-Muller.
TIS_HD double u01_open_closed(uint32_t r) { return ((double)r + 1.0) * 2.3283064365386963e-10; }

// Stream tags of the counter's fourth word.
constexpr uint32_t kTagNoise = 0x4E4F4953u;  // "NOIS"
constexpr uint32_t kTagHb = 0x48425253u;     // "HBRS"
constexpr uint32_t kTagSwap = 0x52454D44u;   // "REMD"

// ------------------------------------------------------------------ forces
// Gradient pieces are written into g[] (one V3 per bead) and scaled by the
// caller, so that a force can be added with one atomic per component.

// Harmonic bond k (r - r0)^2 between a and b: F_a = -t d, F_b = t d with
// d = x_b - x_a, t = 2 k (r0 - r) / r (stacking_bond_force, fast.cpp:957-982).
template <class R>
TIS_DEV R bond_coef(V3<R> d, R k, R r0, R *r_out) {
    R r = sqrt(dot(d, d));
    if (r_out) *r_out = r;
    return (R)2 * k * (r0 - r) / r;
}

// A cosine clamped to [-1, 1]: FP32 rounding carries the cosine of a straight
// (or zero) angle just past -1 (or 1), where acos and sqrt(1 - c^2) are NaN;
// inside [-1, 1] the value is unchanged.
template <class R>
TIS_DEV R unit_clamp(R c) {
    return c < (R)-1 ? (R)-1 : (c > (R)1 ? (R)1 : c);
}

// Angle at b of (a, b, c): returns theta; g[0], g[2] = d cos(theta) / d x_a, x_c.
template <class R>
TIS_DEV R angle_grad(V3<R> ba, V3<R> bc, R *cos_out, V3<R> *g0, V3<R> *g2) {
    R ia = (R)1 / dot(ba, ba), ib = (R)1 / dot(bc, bc);
    R iab = sqrt(ia * ib);
    R c = dot(ba, bc) * iab;
    *cos_out = c;
    *g0 = bc * iab - ba * (c * ia);
    *g2 = ba * iab - bc * (c * ib);
    return c;
}

// Dihedral of (p1, p2, p3, p4) = atan2((b1.n)|b2|, m.n), m = b1 x b2, n = b2 x b3
// (dihedral_angle, fast.cpp:1255-1302), with its gradient by the
// singularity-free cross-product form.
template <class R>
TIS_DEV R dihedral_grad(V3<R> b1, V3<R> b2, V3<R> b3, V3<R> g[4]) {
    V3<R> m = cross(b1, b2), n = cross(b2, b3);
    R b22 = dot(b2, b2);
    R lb2 = sqrt(b22);
    R psi = atan2(dot(b1, n) * lb2, dot(m, n));
    // collinear beads (m or n = 0) leave the angle undefined: no force from it
    const R mm = dot(m, m), nn = dot(n, n);
    V3<R> g1 = mm > (R)0 ? m * (-lb2 / mm) : m * (R)0;
    V3<R> g4 = nn > (R)0 ? n * (lb2 / nn) : n * (R)0;
    R p12 = dot(b1, b2) / b22, p32 = dot(b3, b2) / b22;
    g[0] = g1;
    g[1] = g4 * p32 - g1 * ((R)1 + p12);
    g[2] = g1 * p12 - g4 * ((R)1 + p32);
    g[3] = g4;
    return psi;
}
template <class R>
TIS_DEV R dihedral_only(V3<R> b1, V3<R> b2, V3<R> b3) {
    V3<R> m = cross(b1, b2), n = cross(b2, b3);
    return atan2(dot(b1, n) * sqrt(dot(b2, b2)), dot(m, n));
}

// Target re-wrapped to within pi of the current angle (fast.cpp:1999-2019).
template <class R>
TIS_DEV R wrap_target(R psi, R psi0) {
    const R pi = (R)kPi;
    if (fabs(psi - psi0) <= pi) return psi0;
    if (psi - psi0 > pi) return psi0 + (R)2 * pi;
    return psi0 - (R)2 * pi;
}

}  // namespace tis
