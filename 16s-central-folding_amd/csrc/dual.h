// Forward-mode derivatives for the force-matching backward pass: a value and
// its directional derivative (tangent), carried through the same templated
// force routines that the step uses, so that the tangent of every force is
// d F / d x . v for the position tangent v (a Hessian-vector product, since
// F = -grad U).  FP64 in both parts.
#pragma once
#include "device_math.h"

namespace tis {
// In a namespace of its own so that its math overloads are found by
// argument-dependent lookup only and never hide the plain ones.
namespace ad {

struct Dual {
    double v, d;
    Dual() = default;
    TIS_DEV Dual(double x) : v(x), d(0.0) {}
    TIS_DEV Dual(double x, double dx) : v(x), d(dx) {}
    TIS_DEV explicit operator double() const { return v; }
};

TIS_DEV Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
TIS_DEV Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
TIS_DEV Dual operator-(Dual a) { return {-a.v, -a.d}; }
TIS_DEV Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
TIS_DEV Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
TIS_DEV Dual &operator+=(Dual &a, Dual b) { a = a + b; return a; }
TIS_DEV Dual &operator-=(Dual &a, Dual b) { a = a - b; return a; }
TIS_DEV Dual &operator*=(Dual &a, Dual b) { a = a * b; return a; }
TIS_DEV bool operator<(Dual a, Dual b) { return a.v < b.v; }
TIS_DEV bool operator>(Dual a, Dual b) { return a.v > b.v; }
TIS_DEV bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
TIS_DEV bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
TIS_DEV Dual sqrt(Dual a) {
    const double s = ::sqrt(a.v);
    return {s, a.d / (2.0 * s)};
}
TIS_DEV Dual acos(Dual a) { return {::acos(a.v), -a.d / ::sqrt(1.0 - a.v * a.v)}; }
TIS_DEV Dual atan2(Dual y, Dual x) {
    return {::atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / (x.v * x.v + y.v * y.v)};
}
TIS_DEV Dual exp(Dual a) {
    const double e = ::exp(a.v);
    return {e, e * a.d};
}
TIS_DEV Dual fabs(Dual a) { return a.v < 0.0 ? -a : a; }

}  // namespace ad
using ad::Dual;

// A position with its tangent.
struct DualPt {
    double4 x, t;
};
template <class R>
struct PtOf {
    using type = double4;
};
template <>
struct PtOf<Dual> {
    using type = DualPt;
};

// Difference b - a of two positions with tangents (no periodic image, as diff).
template <class R>
TIS_DEV V3<R> diff(const DualPt &a, const DualPt &b) {
    return {Dual(b.x.x - a.x.x, b.t.x - a.t.x), Dual(b.x.y - a.x.y, b.t.y - a.t.y),
            Dual(b.x.z - a.x.z, b.t.z - a.t.z)};
}

// Value part of a force dotted with the tangent of the point it acts on; 0 for
// plain points.
TIS_DEV double value_dot_tangent(const V3<Dual> &f, const DualPt &p) {
    return f.x.v * p.t.x + f.y.v * p.t.y + f.z.v * p.t.z;
}
template <class R>
TIS_DEV double value_dot_tangent(const V3<R> &, const double4 &) { return 0.0; }

}  // namespace tis
