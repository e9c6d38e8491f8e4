// mcm_physics.h -- per-particle physics of the MCM time step, written once for
// the host (initialisation) and the device (fused gfx950 kernels).
//
// Symmetric 3x3 tensors are held as 6 components in the order
// 11, 22, 33, 12, 13, 23; the spin tensor as its 3 upper components
// W12, W13, W23.  Every expression keeps the evaluation order of the
// reference routine it restates (cited per function), and the library is
// compiled without FP contraction, so for the same inputs these functions
// return the reference's values up to the last-place differences of the
// transcendental functions (tan, pow).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "mcm_dual.h"

#define MCM_HD __host__ __device__ __forceinline__

enum { S11 = 0, S22 = 1, S33 = 2, S12 = 3, S13 = 4, S23 = 5 };

// Material constants (mcm_database.f90:235-257 plus derived values).  The
// scalar type is double in the step; the adjoint instantiates the physics
// below with dual numbers (mcm_dual.h) to take local derivatives.
template <class T>
struct MatParamsT {
    int model, eos, visc_type, ispall, efps_flag;
    int pad_;
    T rho, g, av_l, av_q, h, rho_max;
    T strinput[48];
    T eosinput[24];
    T k_cont, n_cont;   // penalty contact K and n (mcm_contact_in.f90:22-45)
};
typedef MatParamsT<double> MatParams;

// Per-particle state held in registers by the fused kernels.
template <class T>
struct PStateT {
    T mass, rho, rhoold, rho0, h, hold, h0;
    T p, e, etry, einc, c, vabs, pcut, p_cut, efps, fail, q, critts;
    T sig[6], s[6], alfa[6];
    T rod[6], spin[3], tracerod;
};
typedef PStateT<double> PState;

MCM_HD double fsign1(double x) { return signbit(x) ? -1.0 : 1.0; }
MCM_HD double dmax(double a, double b) { return a > b ? a : b; }
MCM_HD double dmin(double a, double b) { return a < b ? a : b; }
MCM_HD double ddim(double a, double b) { return a - b > 0.0 ? a - b : 0.0; }

// Correctly rounded quotients a / b for several a and one divisor b.  On the
// device a / b compiles to a Newton-refined reciprocal of b, the residual
// correction fma(a - b*q, r, q), and scaling/fix-up steps that change
// nothing when |a| and |b| lie in [2^-300, 2^300] (no scaling for those
// exponents, normal quotient).  For such operands the refined reciprocal is
// computed once and shared, which gives every quotient bit for bit; other
// operands (and host code) divide.
struct Divisor {
    double b, r;
    bool fast;
    MCM_HD static bool in_range(double x) {
        const double ax = fabs(x);
        return (int)(ax >= 0x1p-300) & (int)(ax <= 0x1p300);
    }
    MCM_HD explicit Divisor(double b_) : b(b_), r(0.0), fast(false) {
#if defined(__HIP_DEVICE_COMPILE__)
        fast = in_range(b);
        double y = __builtin_amdgcn_rcp(b);
        double e = fma(-b, y, 1.0);
        y = fma(y, e, y);
        e = fma(-b, y, 1.0);
        r = fma(y, e, y);
#endif
    }
    // The residual a - b q is formed as -(b q - a): as exact as fma(-b, q, a)
    // and equal to it but for the sign of a zero residual, which leaves a
    // nonzero quotient unchanged and makes a = -0 give -0, as a / b does for
    // b > 0 (fma(-b, q, a) would give +0 there).
    MCM_HD double quot(double a) const {
        const double q = a * r;
        const double rem = -fma(b, q, -a);
        return fma(rem, r, q);
    }
    MCM_HD double div(double a) const {
#if defined(__HIP_DEVICE_COMPILE__)
        if ((int)fast & (int)in_range(a)) return quot(a);
#endif
        return a / b;
    }
    // three quotients with one range test
    MCM_HD void div3(double a0, double a1, double a2, double q[3]) const {
#if defined(__HIP_DEVICE_COMPILE__)
        // bitwise, not short-circuit: one branch instead of four
        if ((int)fast & (int)in_range(a0) & (int)in_range(a1) & (int)in_range(a2)) {
            q[0] = quot(a0);
            q[1] = quot(a1);
            q[2] = quot(a2);
            return;
        }
#endif
        q[0] = a0 / b;
        q[1] = a1 / b;
        q[2] = a2 / b;
    }
};


// ------------------------------------------------------------------ kernels
// Cubic B-spline, 3-D (mcm_kernel.f90:42-95); r = |xi - xj|
MCM_HD double kernel_w(double rad, double havg, double pi) {
    double const1 = 1.0 / (pi * havg * havg * havg);
    double const2 = const1 / 4.0;
    double z = rad / havg;
    // both branches of mcm_kernel.f90:78-92 are evaluated and one selected
    // (no divergence; each branch's expression as in the reference)
    double w1 = const1 * (((0.75 * z - 1.5) * z) * z + 1.0);
    double t = 2.0 - z;
    double w2 = const2 * (t * (t * t));
    const double w12 = __builtin_unpredictable(z < 2.0) ? w2 : 0.0;
    return __builtin_unpredictable(z < 1.0) ? w1 : w12;
}

// dW/dr factor of mcm_gradw.f90:41-103; gradient_n = dw * xij_n / (rad * havg)
MCM_HD double kernel_dw(double rad, double havg, double pi) {
    double const1 = 1.0 / (pi * havg * havg * havg);
    double const2 = const1 * 0.25;
    double z = rad / havg;
    double d1 = const1 * ((2.25 * z - 3.0) * z);
    double t = 2.0 - z;
    double d2 = -(3.0 * const2 * (t * t));
    const double d12 = __builtin_unpredictable(z < 2.0) ? d2 : 0.0;
    return __builtin_unpredictable(z < 1.0) ? d1 : d12;
}

// gradient of W with respect to xi, xij = xi - xj (mcm_gradw.f90:76-101);
// zero at rad == 0
MCM_HD void kernel_grad(double xij0, double xij1, double xij2, double havg, double pi, double* dwdx,
                        double* rad_out) {
    double rad = 0.0;
    rad = rad + xij0 * xij0;
    rad = rad + xij1 * xij1;
    rad = rad + xij2 * xij2;
    rad = sqrt(rad);
    *rad_out = rad;
    double dw = kernel_dw(rad, havg, pi);
    const Divisor rh(rad * havg);
    double g[3];
    rh.div3(dw * xij0, dw * xij1, dw * xij2, g);
    const bool zero = rad == 0.0;
    dwdx[0] = __builtin_unpredictable(zero) ? 0.0 : g[0];
    dwdx[1] = __builtin_unpredictable(zero) ? 0.0 : g[1];
    dwdx[2] = __builtin_unpredictable(zero) ? 0.0 : g[2];
}

// ---------------------------------------------------- pair loops, fast path
// kernel_grad and kernel_w above test every quotient's operands, and a pair
// whose test fails (the self pair, rad = 0; a ghost with a zero coordinate
// difference) sends the wavefront through full divisions.  The fast path
// below drops the per-pair tests: it is taken for a whole launch when every
// pair record's coordinates are 0 or of magnitude in [2^-30, 2^30] and every
// h lies in [2^-30, 2^30] (fast_coord / fast_h, checked where the records
// are written).  From those ranges, for every pair:
//   xij in {0} u [2^-82, 2^31] (both coordinates multiples of 2^-82), rad^2 in
//   {0} u [2^-164, 2^64] (no scaling step in the square root), rad in {0} u
//   [2^-82, 2^32], pi havg^3 in [2^-89, 2^92], z = rad / havg in {0} u
//   [2^-112, 2^62], dW/dr in {0} u [2^-205, 2^91], its products with xij in
//   {0} u [2^-287, 2^122] and rad havg in {0} u [2^-112, 2^62];
// every quotient then has operands in Divisor's range (or a zero numerator,
// or rad = 0, whose gradient is replaced by 0), so each result equals the
// checked path's bit for bit.  W lies in {0} u [2^-250, 2^89] and each
// gradient component in {0} u [2^-349, 2^234] (used by fast_sums).
// (1 when x is in range, else 0; int so that several tests combine without
// branches)
MCM_HD int fast_coord(double x) {
    const double a = fabs(x);
    return (int)(x == 0.0) | ((int)(a >= 0x1p-30) & (int)(a <= 0x1p30));
}
MCM_HD int fast_h(double h) { return (int)(h >= 0x1p-30) & (int)(h <= 0x1p30); }

#if defined(__HIP_DEVICE_COMPILE__)
// sqrt(x) for x = 0 or x >= 2^-767: the device square root (reciprocal
// square root, then the same Newton steps) without its scaling of tiny
// arguments, which multiplies by 2^0 in that range
__device__ __forceinline__ double sqrt_noscale(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    double e = fma(-g, g, x);
    h = fma(h, r, h);
    g = fma(e, h, g);
    e = fma(-g, g, x);
    g = fma(e, h, g);
    return x == 0.0 ? x : g;
}

// W and grad W of one pair (kernel_w and kernel_grad, one evaluation of
// rad, 1/(pi havg^3) and z), fast path
__device__ __forceinline__ void pair_wgrad_fast(double xij0, double xij1, double xij2, double havg, double pi,
                                                double& w, double dwdx[3]) {
    double rad2 = 0.0;
    rad2 = rad2 + xij0 * xij0;
    rad2 = rad2 + xij1 * xij1;
    rad2 = rad2 + xij2 * xij2;
    const double rad = sqrt_noscale(rad2);
    const double const1 = Divisor(pi * havg * havg * havg).quot(1.0);
    const double const2 = const1 * 0.25;   // const1 / 4.0 in kernel_w: the same value
    const double z = Divisor(havg).quot(rad);
    const double t = 2.0 - z;
    // selects, not branches, as in kernel_w / kernel_dw
    const double d1 = const1 * ((2.25 * z - 3.0) * z);
    const double d2 = -(3.0 * const2 * (t * t));
    const double d12 = __builtin_unpredictable(z < 2.0) ? d2 : 0.0;
    const double dw = __builtin_unpredictable(z < 1.0) ? d1 : d12;
    const double w1 = const1 * (((0.75 * z - 1.5) * z) * z + 1.0);
    const double w2 = const2 * (t * (t * t));
    const double w12 = __builtin_unpredictable(z < 2.0) ? w2 : 0.0;
    w = __builtin_unpredictable(z < 1.0) ? w1 : w12;
    const Divisor rh(rad * havg);
    const double g0 = rh.quot(dw * xij0), g1 = rh.quot(dw * xij1), g2 = rh.quot(dw * xij2);
    const bool zero = rad == 0.0;
    dwdx[0] = __builtin_unpredictable(zero) ? 0.0 : g0;
    dwdx[1] = __builtin_unpredictable(zero) ? 0.0 : g1;
    dwdx[2] = __builtin_unpredictable(zero) ? 0.0 : g2;
}

__device__ __forceinline__ void pair_grad_fast(double xij0, double xij1, double xij2, double havg, double pi,
                                               double dwdx[3]) {
    double w;
    pair_wgrad_fast(xij0, xij1, xij2, havg, pi, w, dwdx);
}

// W and grad W of one pair for the one-pass corrected kernels, which are
// held to a tolerance rather than to the reference's bits: one reciprocal
// of havg and the reciprocal square root of rad^2, each refined by one
// Newton step (relative error near 2^-50), instead of three correctly
// rounded quotients and a square root, with contracted multiply-adds; ipi =
// 1 / pi.  Same operand ranges as pair_wgrad_fast (the fast path's); the
// self pair (rad = 0) gives W(0) and a zero gradient.
__device__ __forceinline__ void pair_wgrad_1p(double xij0, double xij1, double xij2, double havg, double ipi,
                                              double& w, double dwdx[3]) {
#pragma clang fp contract(fast)
    const double rad2 = xij0 * xij0 + xij1 * xij1 + xij2 * xij2;
    double y = __builtin_amdgcn_rsq(rad2);
    y = y * (1.5 - 0.5 * rad2 * (y * y));
    const bool zero = rad2 == 0.0;
    const double rad = zero ? 0.0 : rad2 * y;
    double ih = __builtin_amdgcn_rcp(havg);
    ih = fma(ih, fma(-havg, ih, 1.0), ih);
    const double const1 = ipi * (ih * ih * ih);
    const double const2 = const1 * 0.25;
    const double z = rad * ih;
    const double t = 2.0 - z;
    const double d1 = const1 * ((2.25 * z - 3.0) * z);
    const double d2 = -(3.0 * const2 * (t * t));
    const double d12 = __builtin_unpredictable(z < 2.0) ? d2 : 0.0;
    const double dw = __builtin_unpredictable(z < 1.0) ? d1 : d12;
    const double w1 = const1 * (((0.75 * z - 1.5) * z) * z + 1.0);
    const double w2 = const2 * (t * (t * t));
    const double w12 = __builtin_unpredictable(z < 2.0) ? w2 : 0.0;
    w = __builtin_unpredictable(z < 1.0) ? w1 : w12;
    const double g = zero ? 0.0 : dw * (ih * y);   // dW/dr / (rad havg)
    dwdx[0] = g * xij0;
    dwdx[1] = g * xij1;
    dwdx[2] = g * xij2;
}
#endif

// W and grad W of one pair: the fast path, or the checked functions
template <bool FAST>
__device__ __forceinline__ void pair_wgrad(double xij0, double xij1, double xij2, double havg, double pi, double& w,
                                           double dwdx[3]) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (FAST) {
        pair_wgrad_fast(xij0, xij1, xij2, havg, pi, w, dwdx);
        return;
    }
#endif
    double rad;
    kernel_grad(xij0, xij1, xij2, havg, pi, dwdx, &rad);
    w = kernel_w(rad, havg, pi);
}

template <bool FAST>
__device__ __forceinline__ void pair_grad(double xij0, double xij1, double xij2, double havg, double pi,
                                          double dwdx[3]) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (FAST) {
        pair_grad_fast(xij0, xij1, xij2, havg, pi, dwdx);
        return;
    }
#endif
    double rad;
    kernel_grad(xij0, xij1, xij2, havg, pi, dwdx, &rad);
}

// The corrected passes' quotients (grad W sw - W sg) / sw^2 on the fast path:
// with sw in [2^-100, 2^100] and each sg component 0 or of magnitude in
// [2^-200, 2^200] (tested once per particle), the two products are 0 or of
// magnitude >= 2^-450, so the numerator is 0 or in [2^-502, 2^335] (both
// products are multiples of 2^-502), and sw^2 lies in [2^-200, 2^200]: no
// operand is below 2^-969, the exponents differ by less than 768 and the
// quotient is normal, so the division's scaling and fix-up steps change
// nothing and Divisor::quot gives a / b bit for bit.
MCM_HD bool fast_sums(double sw, const double sg[3]) {
    int ok = (int)(sw >= 0x1p-100) & (int)(sw <= 0x1p100);
    for (int l = 0; l < 3; l++) {
        const double a = fabs(sg[l]);
        ok &= (int)(sg[l] == 0.0) | ((int)(a >= 0x1p-200) & (int)(a <= 0x1p200));
    }
    return ok != 0;
}

// 3x3 cofactor inverse, identity when |det| < 1e-8 (mcm_math_func.f90:50-79).
// m and inv are column-major (element (r,c) at r + 3c).
MCM_HD void inverse3(const double* m, double* inv) {
#define M_(r, c) m[((r) - 1) + 3 * ((c) - 1)]
#define I_(r, c) inv[((r) - 1) + 3 * ((c) - 1)]
    double sd11 = M_(2, 2) * M_(3, 3) - M_(2, 3) * M_(3, 2);
    double sd12 = M_(2, 1) * M_(3, 3) - M_(2, 3) * M_(3, 1);
    double sd13 = M_(2, 1) * M_(3, 2) - M_(2, 2) * M_(3, 1);
    double sd21 = M_(1, 2) * M_(3, 3) - M_(1, 3) * M_(3, 2);
    double sd22 = M_(1, 1) * M_(3, 3) - M_(1, 3) * M_(3, 1);
    double sd23 = M_(1, 1) * M_(3, 2) - M_(1, 2) * M_(3, 1);
    double sd31 = M_(1, 2) * M_(2, 3) - M_(1, 3) * M_(2, 2);
    double sd32 = M_(1, 1) * M_(2, 3) - M_(1, 3) * M_(2, 1);
    double sd33 = M_(1, 1) * M_(2, 2) - M_(1, 2) * M_(2, 1);
    double det = M_(1, 1) * sd11 - M_(1, 2) * sd12 + M_(1, 3) * sd13;
    if (fabs(det) < 1.0e-08) {
        for (int k = 0; k < 9; k++) inv[k] = 0.0;
        I_(1, 1) = 1.0; I_(2, 2) = 1.0; I_(3, 3) = 1.0;
        return;
    }
    double invdet = 1.0 / det;
    I_(1, 1) = sd11 * invdet; I_(1, 2) = -sd21 * invdet; I_(1, 3) = sd31 * invdet;
    I_(2, 1) = -sd12 * invdet; I_(2, 2) = sd22 * invdet; I_(2, 3) = -sd32 * invdet;
    I_(3, 1) = sd13 * invdet; I_(3, 2) = -sd23 * invdet; I_(3, 3) = sd33 * invdet;
#undef M_
#undef I_
}

// rod / spin / trace from a velocity-gradient accumulator g[n][m]
// (mcm_oldstrain.f90:56-63)
MCM_HD void rod_spin(const double g[3][3], PState& s) {
    s.rod[S11] = 0.5 * (g[0][0] + g[0][0]);
    s.rod[S22] = 0.5 * (g[1][1] + g[1][1]);
    s.rod[S33] = 0.5 * (g[2][2] + g[2][2]);
    s.rod[S12] = 0.5 * (g[0][1] + g[1][0]);
    s.rod[S13] = 0.5 * (g[0][2] + g[2][0]);
    s.rod[S23] = 0.5 * (g[1][2] + g[2][1]);
    s.spin[0] = 0.5 * (g[0][1] - g[1][0]);
    s.spin[1] = 0.5 * (g[0][2] - g[2][0]);
    s.spin[2] = 0.5 * (g[1][2] - g[2][1]);
    s.tracerod = s.rod[S11] + s.rod[S22] + s.rod[S33];
}

// --------------------------------------------------------------- update_h
// mcm_update_h.f90:25-54; returns true when h was updated (enters hmax)
MCM_HD bool update_h(PState& s, int h_opt, double dt) {
    if (h_opt == 1) {
        if (s.p >= s.pcut) {
            s.hold = s.h;
            s.h = s.hold * (1.0 + s.tracerod * dt / 3.0);
            return true;
        }
    } else if (h_opt == 2) {
        if (s.p >= s.pcut) {
            s.hold = s.h;
            s.h = s.h0 * pow(s.rho0 / s.rho, 1.0 / 3.0);
            return true;
        }
    }
    return false;
}

// ----------------------------------------------------------------- rhoupd
// mcm_rhoupd.f90:24-71 (the out-of-range warnings are not reproduced)
template <class T, class M>
MCM_HD void rhoupd(PStateT<T>& s, const M& m, T dt) {
    const double othird = 1.0 / 3.0;
    s.rhoold = s.rho;
    s.rho = s.rho * (1.0 - s.tracerod * dt);
    for (int pass = 0; pass < 2; pass++) {
        if (s.rho > m.rho_max) {
            s.rho = s.rho / (1.0 - s.tracerod * dt);
            s.rod[S11] = s.rod[S11] - othird * s.tracerod;
            s.rod[S22] = s.rod[S22] - othird * s.tracerod;
            s.rod[S33] = s.rod[S33] - othird * s.tracerod;
            s.tracerod = 0.0;
            if (pass == 1) s.h = s.hold;
            s.h = m.h * pow(m.rho / s.rho, 1.0 / 3.0);
        }
    }
}

// ------------------------------------------------------------ stress_rot1
// mcm_stress_rot1.f90:22-184 (iorder = 3); the qq accumulation multiplies a
// zero matrix in the reference and is not carried.
template <class T>
MCM_HD void stress_rot1(PStateT<T>& s, T dt) {
    const double etol = (double)1.0e-3f;   // single-precision literal of the reference
    const double z1 = 1.0 / 3.0, z2 = 2.0 / 15.0, z3 = 17.0 / 315.0, z4 = 62.0 / 2835.0;
    T s11 = s.sig[S11], s22 = s.sig[S22], s33 = s.sig[S33];
    T s12 = s.sig[S12], s23 = s.sig[S23], s31 = s.sig[S13];
    T dt2 = 0.5 * dt;
    T wzz = -(dt2 * s.spin[0]);
    T wyy = dt2 * s.spin[1];
    T wxx = -(dt2 * s.spin[2]);
    T wxx2 = wxx * wxx, wyy2 = wyy * wyy, wzz2 = wzz * wzz;
    T qm2 = wxx2 + wyy2 + wzz2;
    T qm = sqrt(qm2);
    T dd = tan(qm * 0.5);
    T aa, cc;
    if (qm < etol) {
        aa = qm2 * 0.25;
        cc = 0.5 * (1.0 + aa * (z1 + aa * (z2 + aa * (z3 + aa * z4))));
    } else {
        cc = dd / qm;
    }
    T bb = 2.0 * cc / (1.0 + dd * dd);
    aa = bb * cc;
    T Q00 = 1.0 - aa * (wyy2 + wzz2);
    T Q11 = 1.0 - aa * (wxx2 + wzz2);
    T Q22 = 1.0 - aa * (wxx2 + wyy2);
    T Q01 = aa * wxx * wyy;
    T Q02 = aa * wxx * wzz;
    T Q12 = aa * wyy * wzz;
    T Q10 = Q01, Q20 = Q02, Q21 = Q12;
    T bwxx = bb * wxx, bwyy = bb * wyy, bwzz = bb * wzz;
    Q01 = Q01 - bwzz; Q10 = Q10 + bwzz;
    Q02 = Q02 + bwyy; Q20 = Q20 - bwyy;
    Q12 = Q12 - bwxx; Q21 = Q21 + bwxx;
    T qs00 = Q00 * s11 + Q01 * s12 + Q02 * s31;
    T qs01 = Q00 * s12 + Q01 * s22 + Q02 * s23;
    T qs02 = Q00 * s31 + Q01 * s23 + Q02 * s33;
    T qs10 = Q10 * s11 + Q11 * s12 + Q12 * s31;
    T qs11 = Q10 * s12 + Q11 * s22 + Q12 * s23;
    T qs12 = Q10 * s31 + Q11 * s23 + Q12 * s33;
    T qs20 = Q20 * s11 + Q21 * s12 + Q22 * s31;
    T qs21 = Q20 * s12 + Q21 * s22 + Q22 * s23;
    T qs22 = Q20 * s31 + Q21 * s23 + Q22 * s33;
    s.sig[S11] = qs00 * Q00 + qs01 * Q01 + qs02 * Q02;
    s.sig[S22] = qs10 * Q10 + qs11 * Q11 + qs12 * Q12;
    s.sig[S33] = qs20 * Q20 + qs21 * Q21 + qs22 * Q22;
    s.sig[S12] = qs00 * Q10 + qs01 * Q11 + qs02 * Q12;
    s.sig[S23] = qs10 * Q20 + qs11 * Q21 + qs12 * Q22;
    s.sig[S13] = qs00 * Q20 + qs01 * Q21 + qs02 * Q22;
}

// energy increment before the stress update (mcm_f3dm3.f90:54-56); q is isotropic
template <class T>
MCM_HD T einc_pre(const PStateT<T>& s, T dt) {
    const T* r = s.rod;
    const T* g = s.sig;
    return (r[S11] * (g[S11] + s.q) + r[S22] * (g[S22] + s.q) + r[S33] * (g[S33] + s.q) +
            2.0 * (r[S12] * (g[S12] + 0.0) + r[S13] * (g[S13] + 0.0) + r[S23] * (g[S23] + 0.0))) *
           dt;
}

// energy increment after the stress update (mcm_f3dm3.f90:154-156)
template <class T>
MCM_HD T einc_post(const PStateT<T>& s, T dt) {
    const T* r = s.rod;
    const T* g = s.sig;
    return 0.5 * ((r[S11] * g[S11] + r[S22] * g[S22] + r[S33] * g[S33] +
                   2.0 * (r[S12] * g[S12] + r[S13] * g[S13] + r[S23] * g[S23])) *
                      dt +
                  s.einc);
}

#define CM5(a, b) m.strinput[((a) - 1) + 5 * ((b) - 1)]

// mcm_f3dm1.f90:27-71 (elastic)
template <class T, class M>
MCM_HD void f3dm1(PStateT<T>& s, const M& m, T dt) {
    const double othird = 1.0 / 3.0;
    T ym = CM5(1, 1), pr = CM5(1, 2), gg = CM5(5, 4);
    T gdt = dt * gg;
    T gd2 = 2.0 * gdt;
    T blk = dt * ym / (1.0 - 2.0 * pr);
    T davg = othird * s.tracerod;
    T p_incr = blk * davg;
    s.einc = einc_pre(s, dt);
    s.sig[S11] = s.sig[S11] + p_incr + gd2 * (s.rod[S11] - davg);
    s.sig[S22] = s.sig[S22] + p_incr + gd2 * (s.rod[S22] - davg);
    s.sig[S33] = s.sig[S33] + p_incr + gd2 * (s.rod[S33] - davg);
    s.sig[S12] = s.sig[S12] + gd2 * s.rod[S12];
    s.sig[S13] = s.sig[S13] + gd2 * s.rod[S13];
    s.sig[S23] = s.sig[S23] + gd2 * s.rod[S23];
    s.einc = einc_post(s, dt);
    s.p = -(othird * (s.sig[S11] + s.sig[S22] + s.sig[S33]));
}

// mcm_f3dm3.f90:31-156 (elastic-plastic with kinematic/isotropic hardening);
// returns the yield active-set flag (scle)
template <class T, class M>
MCM_HD int f3dm3(PStateT<T>& s, const M& m, T dt) {
    const double othird = 1.0 / 3.0;
    T qh = CM5(2, 4), qb = CM5(1, 5), qs = CM5(1, 3), ym = CM5(1, 1), pr = CM5(1, 2);
    T efps_f = CM5(2, 6);
    T gg = 2.0 * CM5(5, 4);
    T gdt = dt * gg;
    T qbqh = qb * qh;
    T blk = dt * ym / (1.0 - 2.0 * pr);
    T davg = othird * s.tracerod;
    T p_incr = blk * davg;
    s.p = s.p - p_incr;
    s.einc = einc_pre(s, dt);
    s.sig[S11] = s.sig[S11] + p_incr + gdt * (s.rod[S11] - davg);
    s.sig[S22] = s.sig[S22] + p_incr + gdt * (s.rod[S22] - davg);
    s.sig[S33] = s.sig[S33] + p_incr + gdt * (s.rod[S33] - davg);
    s.sig[S12] = s.sig[S12] + gdt * s.rod[S12];
    s.sig[S13] = s.sig[S13] + gdt * s.rod[S13];
    s.sig[S23] = s.sig[S23] + gdt * s.rod[S23];
    T k = qs + qbqh * s.efps;
    T d11 = s.p + s.sig[S11] - s.alfa[S11];
    T d22 = s.p + s.sig[S22] - s.alfa[S22];
    T d33 = s.p + s.sig[S33] - s.alfa[S33];
    T d12 = s.sig[S12] - s.alfa[S12];
    T d13 = s.sig[S13] - s.alfa[S13];
    T d23 = s.sig[S23] - s.alfa[S23];
    s.s[S11] = d11; s.s[S22] = d22; s.s[S33] = d33; s.s[S12] = d12; s.s[S13] = d13; s.s[S23] = d23;
    T j2 = d23 * d23 + d13 * d13 + d12 * d12 - d11 * d22 - d22 * d33 - d11 * d33;
    T k2 = 3.0 * j2 - k * k;
    T scle = 0.5 * (1.0 + fsign1(k2));
    T fac1 = 1.0 / (1.5 * gg + qh);
    T fac2 = 1.5 * gg;
    T fac3 = (1.0 - qb) * qh;
    T j1 = sqrt(3.0 * fabs(j2)) + 1.0 - scle;
    T defps = scle * fac1 * (j1 - k);
    s.efps = s.efps + defps;
    T des = scle * fac2 * defps / j1;
    T depn = scle * fac3 * defps / j1;
    s.sig[S11] = s.sig[S11] - des * d11;
    s.sig[S22] = s.sig[S22] - des * d22;
    s.sig[S33] = s.sig[S33] - des * d33;
    s.sig[S12] = s.sig[S12] - des * d12;
    s.sig[S13] = s.sig[S13] - des * d13;
    s.sig[S23] = s.sig[S23] - des * d23;
    s.alfa[S11] = s.alfa[S11] + depn * d11;
    s.alfa[S22] = s.alfa[S22] + depn * d22;
    s.alfa[S33] = s.alfa[S33] + depn * d33;
    s.alfa[S12] = s.alfa[S12] + depn * d12;
    s.alfa[S13] = s.alfa[S13] + depn * d13;
    s.alfa[S23] = s.alfa[S23] + depn * d23;
    if (m.efps_flag == 1 && s.efps > efps_f) {
        T press = othird * (s.sig[S11] + s.sig[S22] + s.sig[S33]);
        s.sig[S11] = press; s.sig[S22] = press; s.sig[S33] = press;
        s.sig[S12] = 0.0; s.sig[S13] = 0.0; s.sig[S23] = 0.0;
    }
    s.einc = einc_post(s, dt);
    return scle > 0.5 ? 1 : 0;
}

// mcm_f3dm9.f90:26-48 (viscous fluid).  The reference indexes cm(1:2, mat)
// by the material id and never writes s(1,2) (s(1,3) is written twice).
MCM_HD void f3dm9(PState& s, const MatParams& m, int matid1) {
    const double othird = 1.0 / 3.0;
    double xmu = m.strinput[1 + 5 * (matid1 - 1)];
    if (xmu == 0.0) {
        for (int k = 0; k < 6; k++) s.s[k] = 0.0;
    } else {
        double davg = -(othird * s.tracerod);
        s.s[S11] = xmu * (s.rod[S11] + davg);
        s.s[S22] = xmu * (s.rod[S22] + davg);
        s.s[S33] = xmu * (s.rod[S33] + davg);
        s.s[S13] = xmu * s.rod[S12];
        s.s[S13] = xmu * s.rod[S13];
        s.s[S23] = xmu * s.rod[S23];
    }
}

// mcm_f3dm10.f90:28-157 (elastic-plastic-hydrodynamic).  defps_prev stands
// in for the plastic-strain increment the reference reads uninitialised when
// a particle has failed (:82-89); returns the yield active-set flag.
template <class T, class M>
MCM_HD int f3dm10(PStateT<T>& s, const M& m, T dt, double& defps_prev) {
    const double othird = 1.0 / 3.0;
    const auto* cm = m.strinput;
    T gg = cm[0], qs = cm[1], qh = cm[2], a1 = cm[4], a2 = cm[5];
    int ispall = m.ispall;
    T epf = cm[7];
    T g2 = 2.0 * dt * gg;
    T davg = othird * s.tracerod;
    s.p = -(othird * (s.sig[S11] + s.sig[S22] + s.sig[S33]));
    s.p_cut = s.pcut;
    T ak = qs + qh * s.efps + (a1 + a2 * s.p) * dmax(0.0, s.p);
    s.s[S11] = s.sig[S11] + s.p + g2 * (s.rod[S11] + davg);
    s.s[S22] = s.sig[S22] + s.p + g2 * (s.rod[S22] + davg);
    s.s[S33] = s.sig[S33] + s.p + g2 * (s.rod[S33] + davg);
    s.s[S12] = s.sig[S12] + 0.5 * g2 * s.rod[S12];
    s.s[S13] = s.sig[S13] + 0.5 * g2 * s.rod[S13];
    s.s[S23] = s.sig[S23] + 0.5 * g2 * s.rod[S23];
    if (ispall == 3) {
        T deltap = s.p - s.p_cut;
        T scle = 0.5 * (1.0 + fsign1(deltap));
        s.pcut = scle * s.p_cut;
        for (int k = 0; k < 6; k++) s.s[k] = scle * s.s[k];
        s.p_cut = s.pcut;
    }
    T* d = s.s;
    T aj2 = 0.5 * (d[S11] * d[S11] + d[S22] * d[S22] + d[S33] * d[S33]) + d[S12] * d[S12] +
                 d[S13] * d[S13] + d[S23] * d[S23];
    T sj2 = sqrt(3.0 * aj2);
    T defps = defps_prev;
    int active = 0;
    if (s.fail == 1.0) {
        T thrg = othird / gg;
        defps = ddim(sj2, ak) / ((double)1.e-30f + ddim(1.0, -(thrg * qh)));
        defps = thrg * dmin(sj2, defps);
        s.efps = s.efps + defps;
        active = sj2 > ak ? 1 : 0;
    }
    defps_prev = dval(defps);
    ak = dmax(ak + qh * defps, 0.0);
    T seps = (double)1.e-30f * gg;
    T scle = (ak + seps) / (dmax(ak, sj2) + seps);
    for (int k = 0; k < 6; k++) d[k] = scle * d[k];
    if (epf != 0.0 && s.efps > epf) {
        s.fail = 0.0;
        for (int k = 0; k < 6; k++) d[k] = 0.0;
    }
    if (ispall == 2) {
        aj2 = 0.5 * (d[S11] * d[S11] + d[S22] * d[S22] + d[S33] * d[S33]) + d[S12] * d[S12] + d[S13] * d[S13] +
              d[S23] * d[S23] + (double)1.e-12f;
        sj2 = d[S11] * d[S23] * d[S23] + d[S22] * d[S13] * d[S13] + d[S33] * d[S12] * d[S12] -
              d[S11] * d[S22] * d[S33] - 2.0 * d[S12] * d[S13] * d[S23];
        T akt = -(sqrt(27.0 / aj2) * sj2 * 0.5 / aj2);
        akt = copysign(dmin(fabs(akt), 1.0), akt);
        T ywh = acos(akt) * othird;
        sj2 = 2.0 * sqrt(aj2 * othird) * cos(ywh);
        T davg1 = s.p - sj2 - s.p_cut;
        T sc2 = 0.5 * (1.0 + fsign1(davg1));
        s.pcut = sc2 * s.p_cut;
        for (int k = 0; k < 6; k++) d[k] = sc2 * d[k];
        s.p_cut = s.pcut;
    }
    return active;
}

// mcm_nodalav.f90:32-47 (isotropic nodal viscosity; visc_type 1 only)
template <class T, class M>
MCM_HD void nodalav(PStateT<T>& s, const M& m) {
    s.q = 0.0;
    if (s.tracerod < 0.0) {
        T ht = s.h * s.tracerod;
        s.q = m.av_l * s.rho * s.c * s.h * fabs(s.tracerod) + m.av_q * s.rho * (ht * ht);
    }
}

// mcm_lieupd.f90:26-33 (q:D over the column-major flattening; q is diagonal)
template <class T>
MCM_HD void lieupd(PStateT<T>& s, T dt) {
    T qpm = 0.0;
    qpm = qpm + s.q * s.rod[S11];
    qpm = qpm + s.q * s.rod[S22];
    qpm = qpm + s.q * s.rod[S33];
    s.e = s.e + (s.einc - dt * qpm) * s.mass / s.rho;
}

// mcm_hieupd.f90:28-64 (trial energy)
template <class T>
MCM_HD void hieupd(PStateT<T>& s, T dt) {
    const double othird = 1.0 / 3.0;
    T pold = -(othird * (s.sig[S11] + s.sig[S22] + s.sig[S33]));
    T trace_q = othird * (s.q + s.q + s.q);
    T volold = s.mass / s.rhoold;
    T volnew = s.mass / s.rho;
    // column-major sum over (k,l): 11 21 31 12 22 32 13 23 33
    T e11 = s.rod[S11] * (0.5 * (s.sig[S11] + pold + s.s[S11]) - s.q + trace_q);
    T e22 = s.rod[S22] * (0.5 * (s.sig[S22] + pold + s.s[S22]) - s.q + trace_q);
    T e33 = s.rod[S33] * (0.5 * (s.sig[S33] + pold + s.s[S33]) - s.q + trace_q);
    T e12 = s.rod[S12] * (0.5 * (s.sig[S12] + s.s[S12]) - 0.0);
    T e13 = s.rod[S13] * (0.5 * (s.sig[S13] + s.s[S13]) - 0.0);
    T e23 = s.rod[S23] * (0.5 * (s.sig[S23] + s.s[S23]) - 0.0);
    T vavg = 0.5 * (volnew + volold);
    T eincr = 0.0;
    eincr = eincr + e11; eincr = eincr + e12; eincr = eincr + e13;
    eincr = eincr + e12; eincr = eincr + e22; eincr = eincr + e23;
    eincr = eincr + e13; eincr = eincr + e23; eincr = eincr + e33;
    eincr = vavg * eincr;
    T dvol = volnew - volold;
    T mean_incr = dvol * (0.5 * pold + trace_q);
    s.etry = s.e + dt * eincr - mean_incr;
}

// Sound speed (mcm_sueos.f90:24-46 -> mcm_eos{1,4,13,28,41}c.f90); returns
// nonzero when eos41 meets a negative c^2 (the reference stops there)
template <class T, class M>
MCM_HD int sueos(PStateT<T>& s, const M& m) {
    const auto* ei = m.eosinput;
    T rho = s.rho, rho0 = s.rho0, mass = s.mass;
    switch (m.eos) {
        case 1: {
            T c1 = ei[1], c2 = ei[2], c3 = ei[3], c4 = ei[4], c5 = ei[5], c6 = ei[6];
            T mu = rho / rho0 - 1.0;
            T mu_t = dmax(mu, 0.0);
            T e1 = 0.0;
            T dmudrho = 1.0 / rho0;
            T termone = c1 * dmudrho + (c2 * (dmudrho * dmudrho)) + (3.0 * c3 * dmudrho) * mu * mu;
            T termtwo = (c5 * dmudrho + c6 * dmudrho * dmudrho) * e1;
            T dpdrho = termone + termtwo;
            T dpde = c4 + c5 * mu + c6 * (mu_t * mu_t);
            T mr = mass / rho;
            T c_sq = ((4.0 / 3.0) * m.g / rho0 + dpdrho + s.p / rho0 * (mr * mr) * dpde);
            s.c = c_sq > 0.0 ? sqrt(c_sq) : sqrt(fabs(c_sq));
            return 0;
        }
        case 4: {
            T c1 = ei[0], s1 = ei[1], s2 = ei[2], s3 = ei[3], gamma = ei[4], a1 = ei[5];
            T e1 = s.e * rho0 / mass;
            T mu = rho / rho0 - 1.0;
            T dpdrho;
            if (mu > 0.0) {
                T termone = mu * (1.0 - (gamma / 2.0));
                T termtwo = (a1 / 2.0) * mu * mu;
                T eosnum = mu * rho0 * (c1 * c1) * (1.0 + termone - termtwo);
                T mp1 = mu + 1.0;
                T mu11 = mu / mp1;
                T mu21 = (mu * mu) / mp1;
                T mu22 = (mu * mu) / (mp1 * mp1);
                T mu32 = (mu * mu * mu) / (mp1 * mp1);
                T mu33 = (mu * mu * mu) / (mp1 * mp1 * mp1);
                T eosden = 1.0 - (s1 - 1.0) * mu - s2 * mu21 - s3 * mu32;
                T termthree = rho0 * (c1 * c1) * (1.0 + 2.0 * termone - 3.0 * termtwo);
                T termfour = 1.0 - s1 - 2.0 * s2 * mu11 + s2 * mu22 - 3.0 * s3 * mu22 + 2.0 * s3 * mu33;
                T ed2 = eosden * eosden;
                T ed3 = eosden * ed2;
                dpdrho = (termthree / ed2 - 2.0 * eosnum * termfour / ed3) / rho0 + a1 * e1 / rho0;
            } else {
                dpdrho = c1 * c1 + a1 * e1 / rho0;
            }
            T dpde = gamma + a1 * mu;
            T mr = mass / rho;
            T c_sq = ((4.0 / 3.0) * m.g / rho0 + dpdrho + s.p / rho0 * (mr * mr) * dpde);
            s.c = c_sq > 0.0 ? sqrt(c_sq) : sqrt(fabs(c_sq));
            return 0;
        }
        case 13: {
            T gamma = ei[0];
            T e1 = s.e / mass;
            s.c = sqrt((gamma - 1.0) * gamma * e1);
            return 0;
        }
        case 28: s.c = sqrt(ei[0] * ei[1] / rho0); return 0;
        case 41: {
            T cc = ei[0], sl = ei[1], gamma0 = ei[4];
            T e1 = s.e / mass;
            T ao = rho0 * (cc * cc);
            T mu = rho / rho0 - 1.0;
            T gamma = gamma0 * rho0 / rho;
            T bo = ao * (1.0 + 2.0 * (sl - 1.0));
            T slm = sl - 1.0;
            T co = ao * (2.0 * (sl - 1.0) + 3.0 * (slm * slm));
            T f = gamma * mu;
            T a1 = ao * (1.0 - f);
            T b1 = bo * (2.0 - 1.5 * f);
            T c1 = co * (3.0 - 2.0 * f);
            T dpdrho, c2;
            if (mu > 0.0) dpdrho = (a1 + b1 * mu + c1 * (mu * mu)) / rho0 + gamma * e1;
            else dpdrho = ao * (1.0 - f) / rho0 + gamma * e1;
            if (s.p == s.pcut) {
                c2 = cc * cc;
            } else {
                T dpdi = gamma * rho;
                T didrho = s.p / (rho * rho);
                dpdrho = dpdrho + dpdi * didrho;
                c2 = dpdrho;
            }
            s.c = sqrt(c2);
            return c2 < 0.0 ? 1 : 0;
        }
        default: return 0;
    }
}

// Pressure and energy (mcm_eqos.f90:24-55 -> mcm_eqos{1,4,13,28,41}.f90)
template <class T, class M>
MCM_HD void eqos(PStateT<T>& s, const M& m, T dt) {
    const auto* ei = m.eosinput;
    T rho = s.rho, rho0 = s.rho0, mass = s.mass;
    switch (m.eos) {
        case 1: {
            T c0 = ei[0], c1 = ei[1], c2 = ei[2], c3 = ei[3], c4 = ei[4], c5 = ei[5], c6 = ei[6];
            T mu = rho / rho0 - 1.0;
            T mu_t = dmax(mu, 0.0);
            T dvol = mass / rho - mass / s.rhoold;
            T vol0 = mass / rho0;
            T e1try = s.etry / vol0;
            T termone = c0 + (c1 + c3 * (mu * mu)) * mu + c2 * (mu_t * mu_t);
            T termtwo = c4 + c5 * mu + c6 * (mu_t * mu_t);
            s.p = (termone + termtwo * e1try) / (1.0 + termtwo * (dvol / vol0));
            if (s.p < s.pcut) {
                s.p = s.pcut;
                s.rho = s.rho / (1.0 - s.tracerod * dt);
            } else {
                s.e = s.etry - 0.5 * dvol * s.p;
            }
            break;
        }
        case 4: {
            T c1 = ei[0], s1 = ei[1], s2 = ei[2], s3 = ei[3], gamma = ei[4], a1 = ei[5];
            T mu = rho / rho0 - 1.0;
            T dvol = mass / rho - mass / s.rhoold;
            T vol0 = mass / rho0;
            T e1try = s.etry / vol0;
            if (mu > 0.0) {
                T termone = mu * (1.0 - (gamma / 2.0));
                T termtwo = (a1 / 2.0) * mu * mu;
                T eosnum = mu * rho0 * (c1 * c1) * (1.0 + termone - termtwo);
                T mp1 = mu + 1.0;
                T mu21 = (mu * mu) / mp1;
                T mu32 = (mu * mu * mu) / (mp1 * mp1);
                T eosden = 1.0 - (s1 - 1.0) * mu - s2 * mu21 - s3 * mu32;
                s.p = (eosnum / (eosden * eosden) + (gamma + a1 * mu) * e1try) /
                      (1.0 + 0.5 * (gamma + a1 * mu) * dvol / vol0);
            } else {
                s.p = (rho0 * (c1 * c1) * mu + (gamma + a1 * mu) * e1try) /
                      (1.0 + 0.5 * (gamma + a1 * mu) * dvol / vol0);
            }
            if (s.p < s.pcut) {
                s.p = s.pcut;
                s.rho = s.rho / (1.0 - s.tracerod * dt);
            } else {
                s.e = s.etry - 0.5 * dvol * s.p;
            }
            break;
        }
        case 13: {
            T gamma_m1 = ei[0] - 1.0;
            T e1try = s.etry / mass;
            T dvol = mass / rho - mass / s.rhoold;
            s.p = (rho * gamma_m1 * e1try) / (1.0 + 0.5 * rho * gamma_m1 * dvol / mass);
            s.e = s.etry - 0.5 * dvol * s.p;
            break;
        }
        case 28: {
            T b = ei[0], gamma = ei[1];
            T dvol = mass / rho - mass / s.rhoold;
            s.p = b * (pow(rho / rho0, gamma) - 1.0);
            s.e = s.etry - 0.5 * dvol * s.p;
            break;
        }
        case 41: {
            T cc = ei[0], sl = ei[1], gama0 = ei[4];
            T etry1 = s.etry / mass;
            T ao = rho0 * (cc * cc);
            T mu = rho / rho0 - 1.0;
            T dvol = mass / rho - mass / s.rhoold;
            T vol0 = mass / rho0;
            T gama = gama0 * rho0 / rho;
            T slm = sl - 1.0;
            T bo = ao * (1.0 + 2.0 * (sl - 1.0));
            T co = ao * (2.0 * (sl - 1.0) + 3.0 * (slm * slm));
            T f = gama * mu;
            T ph;
            if (mu > 0.0) ph = ao * mu + bo * (mu * mu) + co * (mu * (mu * mu));
            else ph = ao * mu;
            T pn = (ph * (1.0 - 0.5 * f) + rho * gama * etry1) / (1.0 + 0.5 * rho * gama * dvol / vol0);
            if (pn < s.pcut) {
                pn = s.pcut;
                s.rho = s.rho / (1.0 - s.tracerod * dt);
            } else {
                s.e = etry1 * mass - 0.5 * dvol * pn;
            }
            s.p = pn;
            break;
        }
        default: break;
    }
}

// mcm_stress_up.f90:28-36
template <class T>
MCM_HD void stress_up(PStateT<T>& s) {
    s.sig[S11] = s.s[S11] - s.p;
    s.sig[S22] = s.s[S22] - s.p;
    s.sig[S33] = s.s[S33] - s.p;
    s.sig[S12] = s.s[S12];
    s.sig[S13] = s.s[S13];
    s.sig[S23] = s.s[S23];
}

// per-particle critical timestep (mcm_calccritts.f90:40-78)
template <class T, class M>
MCM_HD T critts(const PStateT<T>& s, const M& m, int tcrit_opt, double mindist) {
    if (tcrit_opt == 0) {
        T mr = s.mass / s.rho;
        T veae = pow(mr, 1.0 / 3.0);
        T ae = pow(mr, 2.0 / 3.0);
        T qu = (m.av_l * s.c + m.av_q * s.mass / s.rho * s.tracerod) / ae;
        return veae / (qu + sqrt(qu * qu + s.c * s.c));
    }
    if (tcrit_opt == 1) return s.h / (s.c + s.vabs);
    return mindist / (s.c + s.vabs);
}

// Material update of one particle in the order of mcm_constitutive.f90:29-83.
// Returns the yield active-set flag of the plastic models.
// MODELS: 0 for any material model, else the bit set (1 << model) of the
// models the run's materials use, which leaves the others out of the code
template <int MODELS = 0>
MCM_HD int constitutive_one(PState& s, const MatParams& m, int matid1, double dt, double& defps_prev,
                            int* eos_err) {
    constexpr bool m1 = MODELS == 0 || (MODELS & (1 << 1)), m3 = MODELS == 0 || (MODELS & (1 << 3));
    constexpr bool m9 = MODELS == 0 || (MODELS & (1 << 9)), m10 = MODELS == 0 || (MODELS & (1 << 10));
    int active = 0;
    switch (m.model) {
        case 1:
            if (m1) { stress_rot1(s, dt); f3dm1(s, m, dt); nodalav(s, m); lieupd(s, dt); }
            break;
        case 3:
            if (m3) { stress_rot1(s, dt); active = f3dm3(s, m, dt); nodalav(s, m); lieupd(s, dt); }
            break;
        case 9:
            if (m9) {
                stress_rot1(s, dt); f3dm9(s, m, matid1);
                *eos_err |= sueos(s, m);
                nodalav(s, m); hieupd(s, dt); eqos(s, m, dt); stress_up(s);
            }
            break;
        case 10:
            if (m10) {
                stress_rot1(s, dt); active = f3dm10(s, m, dt, defps_prev);
                *eos_err |= sueos(s, m);
                nodalav(s, m); hieupd(s, dt); eqos(s, m, dt); stress_up(s);
            }
            break;
        default: break;
    }
    return active;
}

// eos{1,4,13,28,41}p.f90 (initial pressure), used by the initialisation only
template <class T, class M>
MCM_HD void eos_p(PStateT<T>& s, const M& m) {
    const auto* ei = m.eosinput;
    T rho = s.rho, rho0 = s.rho0;
    switch (m.eos) {
        case 1: {
            T mu = rho / rho0 - 1.0, mu_t = dmax(mu, 0.0), e1 = 0.0;
            s.p = ei[0] + (ei[1] + ei[3] * (mu * mu)) * mu + ei[2] * (mu_t * mu_t) +
                  (ei[4] + ei[5] * mu + ei[6] * (mu_t * mu_t)) * e1;
            break;
        }
        case 4: {
            T c1 = ei[0], s1 = ei[1], s2 = ei[2], s3 = ei[3], gamma = ei[4], a1 = ei[5];
            T e1 = s.e * rho0 / s.mass;
            T mu = rho / rho0 - 1.0;
            if (mu > 0.0) {
                T termone = mu * (1.0 - (gamma / 2.0));
                T termtwo = (a1 / 2.0) * mu * mu;
                T eosnum = mu * rho0 * (c1 * c1) * (1.0 + termone - termtwo);
                T mp1 = mu + 1.0;
                T mu21 = (mu * mu) / mp1;
                T mu32 = (mu * mu * mu) / (mp1 * mp1);
                T eosden = 1.0 - (s1 - 1.0) * mu - s2 * mu21 - s3 * mu32;
                s.p = eosnum / (eosden * eosden) + (gamma + a1 * mu) * e1;
            } else {
                s.p = rho0 * (c1 * c1) * mu + (gamma + a1 * mu) * e1;
            }
            break;
        }
        case 13: s.p = rho * (ei[0] - 1.0) * (s.e / s.mass); break;
        case 28: s.p = ei[0] * (pow(rho / rho0, ei[1]) - 1.0); break;
        case 41: {
            T cc = ei[0], sl = ei[1], gamma = ei[4];
            T e1 = s.e / s.mass;
            T ao = rho0 * (cc * cc);
            T mu = rho / rho0 - 1.0;
            T slm = sl - 1.0;
            T bo = ao * (1.0 + 2.0 * (sl - 1.0));
            T co = ao * (2.0 * (sl - 1.0) + 3.0 * (slm * slm));
            gamma = gamma * rho0 / rho;
            T f = gamma * mu, ph;
            if (mu > 0.0) ph = ao * mu + bo * (mu * mu) + co * (mu * (mu * mu));
            else ph = ao * mu;
            T pn = ph * (1.0 - 0.5 * f) + rho * gamma * e1;
            if (pn < s.pcut) pn = s.pcut;
            s.p = pn;
            break;
        }
        default: break;
    }
}
#undef CM5

#include "mcm_cap.h"
