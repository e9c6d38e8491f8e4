// mcm_dual.h -- forward-mode dual numbers for the local Jacobians of the
// adjoint step.
//
// Dual<K> carries a value and K directional derivatives.  The value part of
// every operation is the same IEEE operation as the double code it stands
// in for (no contraction, same library calls), so a function instantiated
// with Dual<K> takes the branches the forward step took and its value part
// equals the forward's bit for bit; the derivative parts are the exact
// derivatives of that branch.  Comparisons compare values.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#define MCM_DHD __host__ __device__ __forceinline__

template <int K>
struct Dual {
    double v;
    double d[K];
    MCM_DHD Dual() : v(0.0) {
        for (int k = 0; k < K; k++) d[k] = 0.0;
    }
    MCM_DHD Dual(double x) : v(x) {   // a constant
        for (int k = 0; k < K; k++) d[k] = 0.0;
    }
};

template <int K> MCM_DHD Dual<K> operator-(const Dual<K>& a) {
    Dual<K> r; r.v = -a.v;
    for (int k = 0; k < K; k++) r.d[k] = -a.d[k];
    return r;
}
template <int K> MCM_DHD Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r; r.v = a.v + b.v;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int K> MCM_DHD Dual<K> operator+(const Dual<K>& a, double b) {
    Dual<K> r = a; r.v = a.v + b;
    return r;
}
template <int K> MCM_DHD Dual<K> operator+(double a, const Dual<K>& b) {
    Dual<K> r = b; r.v = a + b.v;
    return r;
}
template <int K> MCM_DHD Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r; r.v = a.v - b.v;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int K> MCM_DHD Dual<K> operator-(const Dual<K>& a, double b) {
    Dual<K> r = a; r.v = a.v - b;
    return r;
}
template <int K> MCM_DHD Dual<K> operator-(double a, const Dual<K>& b) {
    Dual<K> r; r.v = a - b.v;
    for (int k = 0; k < K; k++) r.d[k] = -b.d[k];
    return r;
}
template <int K> MCM_DHD Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r; r.v = a.v * b.v;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int K> MCM_DHD Dual<K> operator*(const Dual<K>& a, double b) {
    Dual<K> r; r.v = a.v * b;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] * b;
    return r;
}
template <int K> MCM_DHD Dual<K> operator*(double a, const Dual<K>& b) {
    Dual<K> r; r.v = a * b.v;
    for (int k = 0; k < K; k++) r.d[k] = a * b.d[k];
    return r;
}
// d(a/b) = (da - (a/b) db) / b
template <int K> MCM_DHD Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r; r.v = a.v / b.v;
    for (int k = 0; k < K; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
template <int K> MCM_DHD Dual<K> operator/(const Dual<K>& a, double b) {
    Dual<K> r; r.v = a.v / b;
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] / b;
    return r;
}
template <int K> MCM_DHD Dual<K> operator/(double a, const Dual<K>& b) {
    Dual<K> r; r.v = a / b.v;
    for (int k = 0; k < K; k++) r.d[k] = -(r.v * b.d[k]) / b.v;
    return r;
}
template <int K> MCM_DHD Dual<K>& operator+=(Dual<K>& a, const Dual<K>& b) { a = a + b; return a; }
template <int K> MCM_DHD Dual<K>& operator-=(Dual<K>& a, const Dual<K>& b) { a = a - b; return a; }
template <int K> MCM_DHD Dual<K>& operator*=(Dual<K>& a, const Dual<K>& b) { a = a * b; return a; }

#define MCM_DUAL_CMP(op)                                                                           \
    template <int K> MCM_DHD bool operator op(const Dual<K>& a, const Dual<K>& b) { return a.v op b.v; } \
    template <int K> MCM_DHD bool operator op(const Dual<K>& a, double b) { return a.v op b; }          \
    template <int K> MCM_DHD bool operator op(double a, const Dual<K>& b) { return a op b.v; }
MCM_DUAL_CMP(<)
MCM_DUAL_CMP(>)
MCM_DUAL_CMP(<=)
MCM_DUAL_CMP(>=)
MCM_DUAL_CMP(==)
MCM_DUAL_CMP(!=)
#undef MCM_DUAL_CMP

// value of a double or a dual
MCM_DHD double dval(double x) { return x; }
template <int K> MCM_DHD double dval(const Dual<K>& x) { return x.v; }

// f(x) with f'(x) = fp
template <int K> MCM_DHD Dual<K> dual_chain(const Dual<K>& x, double fx, double fp) {
    Dual<K> r; r.v = fx;
    for (int k = 0; k < K; k++) r.d[k] = fp * x.d[k];
    return r;
}
// sqrt at 0 (the second invariant of a zero stress, the spin norm of a
// zero velocity gradient) takes derivative 0: there the forward selects the
// branch that does not use the root's direction
template <int K> MCM_DHD Dual<K> sqrt(const Dual<K>& x) {
    const double s = sqrt(x.v);
    return dual_chain(x, s, s > 0.0 ? 0.5 / s : 0.0);
}
template <int K> MCM_DHD Dual<K> fabs(const Dual<K>& x) { return signbit(x.v) ? -x : x; }
template <int K> MCM_DHD Dual<K> tan(const Dual<K>& x) {
    const double t = tan(x.v);
    return dual_chain(x, t, 1.0 + t * t);
}
template <int K> MCM_DHD Dual<K> cos(const Dual<K>& x) { return dual_chain(x, cos(x.v), -sin(x.v)); }
template <int K> MCM_DHD Dual<K> acos(const Dual<K>& x) {
    return dual_chain(x, acos(x.v), -1.0 / sqrt(1.0 - x.v * x.v));
}
// x^e for a constant exponent (x > 0 where it is used)
template <int K> MCM_DHD Dual<K> pow(const Dual<K>& x, double e) {
    const double p = pow(x.v, e);
    return dual_chain(x, p, e * p / x.v);
}
template <int K> MCM_DHD Dual<K> copysign(const Dual<K>& a, const Dual<K>& b) {
    return signbit(a.v) == signbit(b.v) ? a : -a;
}
// x^y with both operands varying (x > 0)
template <int K> MCM_DHD Dual<K> pow(const Dual<K>& x, const Dual<K>& y) {
    Dual<K> r; r.v = pow(x.v, y.v);
    const double dx = y.v * r.v / x.v, dy = r.v * log(x.v);
    for (int k = 0; k < K; k++) r.d[k] = dx * x.d[k] + dy * y.d[k];
    return r;
}

// the selection helpers of the physics code (max, min, positive difference,
// Fortran sign(1.0, x)) for dual operands, with the same comparisons
template <int K> MCM_DHD Dual<K> dmax(const Dual<K>& a, const Dual<K>& b) { return a > b ? a : b; }
template <int K> MCM_DHD Dual<K> dmax(double a, const Dual<K>& b) { return a > b ? Dual<K>(a) : b; }
template <int K> MCM_DHD Dual<K> dmax(const Dual<K>& a, double b) { return a > b ? a : Dual<K>(b); }
template <int K> MCM_DHD Dual<K> dmin(const Dual<K>& a, const Dual<K>& b) { return a < b ? a : b; }
template <int K> MCM_DHD Dual<K> dmin(double a, const Dual<K>& b) { return a < b ? Dual<K>(a) : b; }
template <int K> MCM_DHD Dual<K> dmin(const Dual<K>& a, double b) { return a < b ? a : Dual<K>(b); }
template <int K> MCM_DHD Dual<K> ddim(const Dual<K>& a, const Dual<K>& b) {
    return a - b > 0.0 ? a - b : Dual<K>(0.0);
}
template <int K> MCM_DHD Dual<K> ddim(double a, const Dual<K>& b) { return a - b > 0.0 ? a - b : Dual<K>(0.0); }
template <int K> MCM_DHD Dual<K> ddim(const Dual<K>& a, double b) { return a - b > 0.0 ? a - b : Dual<K>(0.0); }
template <int K> MCM_DHD double fsign1(const Dual<K>& x) { return signbit(x.v) ? -1.0 : 1.0; }
