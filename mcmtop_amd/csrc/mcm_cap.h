// mcm_cap.h -- the cap model of granular material (material model 4,
// MCM/mcm_f3dm25.f90), per particle, for the host set-up and the fused
// constitutive kernel.
//
// The reference's f3dm25a, feit and fcit have no implicit none: every
// local they do not declare is REAL*4 (a-h, o-z) or INTEGER (i-n), the
// declared ones REAL*8.  Each operation below is carried out in the wider
// kind of its two operands and each assignment rounds to its target's kind,
// as Fortran does; the float literals are the reference's single-precision
// literals.  The library is built without FP contraction, and float
// division and square root are correctly rounded, as on the host.
//
// The reference's Debug configuration compiles with /Qsave: locals are
// static and zero before their first assignment.  The local capa that the
// yield test reads (mcm_f3dm25.f90:205-206) is never assigned, so it is 0;
// blk and davg (:182) are never assigned either, and their product is
// unused.  The mode variable is saved too: a call whose mode tests all fail
// keeps the last call's mode.  With ltype 1 (soil) or 2 (rock) and finite
// values exactly one test passes, so runs are limited to those two types.
//
// Single-precision exp: flang calls the C library's expf, whose result is
// not always correctly rounded (about 1e-4 of arguments differ from the
// correctly rounded value).  cap_expf restates that routine's published
// algorithm (the exp2 table form of ARM's optimized-routines, as built with
// FMA): it equals the host's expf bit for bit on every float argument.
#pragma once

// 2^(i/32) as double bit patterns, minus i << 47 (the exponent the index
// supplies back)
__constant__ __device__ static const unsigned long long kCapExp2Tab[32] = {
    0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
    0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
    0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
    0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
    0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
    0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
    0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
    0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull,
};
static const unsigned long long kCapExp2TabHost[32] = {
    0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
    0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
    0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
    0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
    0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
    0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
    0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
    0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull,
};

MCM_HD float cap_expf(float x) {
    unsigned int ux;
    __builtin_memcpy(&ux, &x, 4);
    const unsigned int abstop = (ux >> 20) & 0x7ff;
    if (abstop >= 0x42b) {   // |x| >= 88, inf or nan
        if (ux == 0xff800000u) return 0.0f;            // exp(-inf)
        if (abstop >= 0x7f8) return x + x;             // inf or nan
        if (x > 0x1.62e42ep6f) return __builtin_inff();  // overflow
        if (x < -0x1.9fe368p6f) return 0.0f;           // underflow
    }
    // x N / ln 2 = k + r, N = 32; exp(x) = 2^(k/N) 2^(r/N) with a cubic in r
    const double InvLn2N = 0x1.71547652b82fep+0 * 32;
    const double Shift = 0x1.8p+52;
    const double C0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32, C1 = 0x1.ebfce50fac4f3p-3 / 32 / 32,
                 C2 = 0x1.62e42ff0c52d6p-1 / 32;
    const double xd = (double)x;
    double kd = fma(InvLn2N, xd, Shift);
    unsigned long long ki;
    __builtin_memcpy(&ki, &kd, 8);
    kd -= Shift;
    const double r = fma(InvLn2N, xd, -kd);
#if defined(__HIP_DEVICE_COMPILE__)
    unsigned long long t = kCapExp2Tab[ki % 32];
#else
    unsigned long long t = kCapExp2TabHost[ki % 32];
#endif
    t += ki << 47;
    double s;
    __builtin_memcpy(&s, &t, 8);
    const double z = fma(C0, r, C1);
    const double r2 = r * r;
    double y = fma(C2, r, 1.0);
    y = fma(z, r2, y);
    y = y * s;
    return (float)y;
}

MCM_HD float cap_maxf(float a, float b) { return a > b ? a : b; }
MCM_HD float cap_minf(float a, float b) { return a < b ? a : b; }

// Material constants of the cap model as its routines read them from
// cm(5, *) (strinput); mcm_f3dm25.f90:110-159.  sets4 has put lambda + 2G in
// cm(5,2), the initial cap parameter in cm(3,3) and the failure surface's
// zero in cm(3,4).
struct CapConst {
    double bulk, capint;
    float shear, alpha, theta, gama, beta, r, dDyna, w, z, t, fet, fept;
    int ivec, ltype;
};

template <class M>
MCM_HD CapConst cap_const(const M& m) {
    const auto* cm = m.strinput;
#define C(a, b) cm[((a) - 1) + 5 * ((b) - 1)]
    CapConst k;
    const float geop = 0.0f;
    k.bulk = C(1, 1);
    k.shear = (float)C(1, 2);
    k.alpha = (float)C(1, 3);
    k.theta = (float)C(1, 4);
    k.gama = (float)C(1, 5);
    k.beta = (float)C(1, 6);
    k.r = (float)C(2, 1);
    k.dDyna = (float)C(2, 2);
    k.w = (float)C(2, 3);
    const float falfac = (float)C(4, 1);
    k.alpha = k.alpha - falfac;
    k.z = (float)C(2, 4);
    k.ivec = C(3, 5) != 0.0 ? 0 : 1;
    k.ltype = (int)C(2, 6);
    const float tcut = (float)C(3, 1), fcut = (float)C(3, 4);
    k.t = cap_maxf(fcut, tcut + 3.0f * geop);
    k.fet = k.alpha - k.gama * cap_expf(-(k.beta * k.t)) + k.theta * k.t;
    k.fept = k.gama * k.beta * cap_expf(-(k.beta * k.t)) + k.theta;
    k.capint = C(3, 3);
#undef C
    return k;
}

// xk and dhdk of one return iteration at the cap parameter capa (the first
// two statements of feit / fcit's loop, mcm_f3dm25.f90:947-948)
MCM_HD float cap_xk(const CapConst& k, double capa) {
    return (float)(capa + (double)k.r * ((double)k.alpha - (double)k.gama * exp(-((double)k.beta * capa)) +
                                         (double)k.theta * capa));
}
MCM_HD float cap_dhdk(const CapConst& k, double capa, float xk) {
    return (float)((double)(k.w * k.dDyna * cap_expf(-(k.dDyna * xk))) *
                   (1.0 + (double)k.r * ((double)k.theta + (double)(k.gama * k.beta) * exp(-((double)k.beta * capa)))));
}

// One failure-surface iteration (feit's loop body, mcm_f3dm25.f90:947-976;
// the four-step vector form of mode 10 repeats the same statements)
MCM_HD void cap_fail_step(const CapConst& k, double& capa, double sig[6], float& x1, float& x2d, float s[6]) {
    const float xk = cap_xk(k, capa);
    const float dhdk = cap_dhdk(k, capa, xk);
    const float omega = k.theta + k.gama * k.beta * cap_expf(-(k.beta * x1));
    const float dcd = (float)((double)k.shear + 9.0 * k.bulk * (double)(omega * omega));
    const float f = x2d - (k.alpha - k.gama * cap_expf(-(k.beta * x1)) + k.theta * x1);
    const float delcap = -(3.0f * omega * f / (dcd * dhdk));
    const float delam = -(dhdk * delcap / (3.0f * omega));
    capa = capa + (double)delcap;
    const float fac1 = k.shear / x2d;
    const float fac2 = (float)(3.0 * k.bulk * (double)omega);
    for (int a = 0; a < 3; a++) sig[a] = sig[a] - (double)(delam * (fac1 * s[a] - fac2));
    for (int a = 3; a < 6; a++) sig[a] = sig[a] - (double)(delam * fac1 * s[a]);
    x1 = (float)(sig[0] + sig[1] + sig[2]);
    for (int a = 0; a < 3; a++) s[a] = (float)(sig[a] - (double)(x1 / 3.0f));
    for (int a = 3; a < 6; a++) s[a] = (float)sig[a];
    x2d = sqrtf(0.5f * (s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) + s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
}

// One cap-surface iteration (fcit's loop body, mcm_f3dm25.f90:1004-1044;
// the four-step vector form of mode 11 repeats the same statements).  xk is
// left at the new cap parameter and elcap at the one before the update: the
// convergence test of fcit reads both.
MCM_HD void cap_cap_step(const CapConst& k, double& capa, double sig[6], float& x1, float& x2d, float s[6],
                         float& xk_out, float& elcap_out) {
    const float unit = 1.0f;
    float xk = cap_xk(k, capa);
    const float dhdk = cap_dhdk(k, capa, xk);
    float dfdk;
    if (capa >= 0.0)
        dfdk = (float)(-2.0 * (fabs((double)x1 - capa) / (double)(k.r * k.r) +
                               ((double)k.alpha - (double)k.gama * exp(-((double)k.beta * capa)) + (double)k.theta * capa) *
                                   ((double)k.theta + (double)(k.gama * k.beta) * exp(-((double)k.beta * capa)))));
    else
        dfdk = (float)((double)(-2.0f * xk) *
                       (1.0 + (double)k.r * ((double)k.theta + (double)(k.gama * k.beta) * exp(-((double)k.beta * capa)))) /
                       (double)(k.r * k.r));
    const float elcap = (float)(capa > (double)(0.0f * unit) ? capa : (double)(0.0f * unit));
    const float omega = cap_minf(-2.0f * fabsf(x1 - elcap) / (k.r * k.r), -1.e-15f * unit);
    const float t2 = 2.0f * (x1 - elcap) / (k.r * k.r);
    const float dcd = (float)((double)(4.0f * k.shear * (x2d * x2d)) + 9.0 * k.bulk * (double)(t2 * t2));
    const float f = x2d * x2d + (x1 - elcap) * (x1 - elcap) / (k.r * k.r) - (xk - elcap) * (xk - elcap) / (k.r * k.r);
    const float delcap = -f / (dfdk + dcd * dhdk / (3.0f * omega));
    const float delam = -(dhdk * delcap / (3.0f * omega));
    capa = capa + (double)delcap;
    xk = cap_xk(k, capa);
    const float fac1 = 2.0f * k.shear;
    const float fac2 = (float)(6.0 * k.bulk * (double)fabsf(x1 - elcap) / (double)(k.r * k.r));
    for (int a = 0; a < 3; a++) sig[a] = sig[a] - (double)(delam * (fac1 * s[a] + fac2));
    for (int a = 3; a < 6; a++) sig[a] = sig[a] - (double)(delam * fac1 * s[a]);
    x1 = (float)(sig[0] + sig[1] + sig[2]);
    for (int a = 0; a < 3; a++) s[a] = (float)(sig[a] - (double)(x1 / 3.0f));
    for (int a = 3; a < 6; a++) s[a] = (float)sig[a];
    x2d = sqrtf(0.5f * (s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) + s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    xk_out = xk;
    elcap_out = elcap;
}

// f3dm25a, mcm_f3dm25.f90:34-838, for one particle: sig and rod as the
// reference's sigma and rod upper triangles (the stress already rotated by
// stress_rot1), capa its cap parameter.  Returns the mode (ptype), or -1
// when the iterated return of mode 10 or 11 did not converge in 100
// iterations (the reference stops).  Writes sig (upper triangle), rod and p.
template <class S>
MCM_HD int f3dm25(S& s, const CapConst& k, double dt, double& capa) {
    const float tol = 1.0e-7f;
    const double othird = 1.0 / 3.0;
    const float third = 1.0f / 3.0f;
    const float unit = 1.0f, geop = 0.0f;
    const float capa_l = 0.0f;   // the never-assigned local capa
    capa = capa > k.capint ? capa : k.capint;
    double sig[6] = {-s.sig[S11], -s.sig[S22], -s.sig[S33], -s.sig[S12], -s.sig[S13], -s.sig[S23]};
    const double dd1 = -s.rod[S11], dd2 = -s.rod[S22], dd3 = -s.rod[S33];
    const double dd4 = -(s.rod[S12] / 2.0), dd5 = -(s.rod[S13] / 2.0), dd6 = -(s.rod[S23] / 2.0);
    const double dd = (dd1 + dd2 + dd3);
    const double davgm = (double)(1.0f / 3.0f) * dd;
    const float dt1 = (float)dt;
    const double p = 3.0 * k.bulk * (double)dt1 * davgm;
    const double gdt = dt * (double)k.shear;
    const double gd2 = 2.0 * gdt;
    const float gtr11 = (float)(sig[0] + p + gd2 * (dd1 - davgm));
    const float gtr22 = (float)(sig[1] + p + gd2 * (dd2 - davgm));
    const float gtr33 = (float)(sig[2] + p + gd2 * (dd3 - davgm));
    const float gtr12 = (float)(sig[3] + gd2 * dd4);
    const float gtr23 = (float)(sig[4] + gd2 * dd5);
    const float gtr31 = (float)(sig[5] + gd2 * dd6);
    const float x1tr = gtr11 + gtr22 + gtr33;
    const float s1 = gtr11 - (1.0f / 3.0f) * x1tr, s2 = gtr22 - (1.0f / 3.0f) * x1tr,
                s3 = gtr33 - (1.0f / 3.0f) * x1tr;
    float x2dtr = 0.5f * (s1 * s1 + s2 * s2 + s3 * s3 + 2.0f * (gtr12 * gtr12 + gtr23 * gtr23 + gtr31 * gtr31));
    x2dtr = sqrtf(x2dtr);
    float fe = k.alpha - k.gama * cap_expf(-(k.beta * capa_l)) + k.theta * capa_l;
    const float xkn = capa_l + k.r * fe;
    const float elcapn = (float)(capa > (double)(0.0f * unit) ? capa : (double)(0.0f * unit));
    const float tmises = fabsf(xkn - elcapn) / k.r;
    fe = k.alpha - k.gama * cap_expf(-(k.beta * x1tr)) + k.theta * x1tr;
    const float f1tr = x2dtr - cap_minf(fe, tmises);
    const float tcrit =
        (float)((double)k.t - 9.0 * k.bulk / (double)k.shear * (double)(x2dtr - k.fet) * (double)k.fept);
    fe = k.alpha - k.gama * cap_expf(-(k.beta * elcapn)) + k.theta * elcapn;
    const float fep = k.beta * k.gama * cap_expf(-(k.beta * elcapn)) + k.theta;
    const float fc2tr = ((xkn - elcapn) * (xkn - elcapn) - (x1tr - elcapn) * (x1tr - elcapn)) / (k.r * k.r);
    const float f2tr = x2dtr * x2dtr - fc2tr;
    const float x1crt =
        (float)((double)elcapn - (double)((x2dtr - fe) * 9.0f) * k.bulk * (double)fep / (double)k.shear);
    fe = (float)((double)k.alpha - (double)k.gama * exp(-((double)k.beta * capa)) + (double)k.theta * capa);
    const double ctol = capa - (double)(tol * fe);
    const float capal = (float)((double)(0.0f * unit) < ctol ? (double)(0.0f * unit) : ctol);
    fe = k.alpha - k.gama * cap_expf(-(k.beta * capal)) + k.theta * capal;
    const float xk = capal + k.r * fe;
    const float evpn = k.w * (1.0f - cap_expf(k.dDyna * (k.z - xkn)));
    const float devpb = k.w * (1.0f - cap_expf(k.dDyna * (k.z - xk))) - evpn;
    const double capan = capa;
    const float t = k.t, fet = k.fet;
    const int ltype = k.ltype;
    int imode = 0;
    if (x1tr <= t && x2dtr <= fet && ltype == 1 && capan > 0.0 && devpb >= 0.0f) imode = 2;
    else if (x1tr <= t && x2dtr <= fet && ltype == 1 && capan > 0.0 && devpb < 0.0f) imode = 7;
    else if (x1tr <= t && x2dtr <= fet && (ltype == 2 || capan <= 0.0)) imode = 3;
    else if (x1tr <= t && x2dtr > fet && x1tr <= tcrit && ltype == 1 && capan > 0.0 && devpb >= 0.0f) imode = 4;
    else if (x1tr <= t && x2dtr > fet && x1tr <= tcrit && ltype == 1 && capan > 0.0 && devpb < 0.0f) imode = 6;
    else if (x1tr <= t && x2dtr > fet && x1tr <= tcrit && (ltype == 2 || capan <= 0.0)) imode = 5;
    else if (x1tr <= t && x2dtr > fet && x1tr > tcrit && capan >= 0.0) imode = 10;
    else if (x1tr <= t && x2dtr > fet && x1tr > tcrit && capan < 0.0) imode = 8;
    else if (x1tr > t && x1tr <= elcapn && f1tr <= tol) imode = 1;
    else if (x1tr > t && x1tr <= elcapn && f1tr > tol && x1tr >= x1crt) imode = 9;
    else if (x1tr > t && x1tr <= elcapn && f1tr > tol && x1tr < x1crt && capan >= 0.0) imode = 10;
    else if (x1tr > t && x1tr <= elcapn && f1tr > tol && x1tr < x1crt && capan < 0.0) imode = 8;
    else if (x1tr > t && x1tr > elcapn && f2tr < tol) imode = 1;
    else if (x1tr > t && x1tr > elcapn && f2tr >= tol) imode = 11;
    float x1 = 0.0f, x2d = 0.0f, ratio;
    double pp;
    const float s_tr[3] = {s1, s2, s3};
    const float g_sh[3] = {gtr12, gtr23, gtr31};
    const float gt[6] = {gtr11, gtr22, gtr33, gtr12, gtr23, gtr31};
    switch (imode) {
        case 1: case 2: case 3:   // :263-315 elastic, tension cut-off
            x1 = imode == 1 ? x1tr : t;
            capa = imode == 2 ? 0.0 : capan;
            pp = (double)(third * x1 + geop);
            for (int a = 0; a < 3; a++) sig[a] = (double)s_tr[a] + pp;
            for (int a = 0; a < 3; a++) sig[3 + a] = (double)g_sh[a];
            break;
        case 4: case 5:   // :320-354 tension corner
            x1 = t;
            x2d = fet;
            capa = imode == 4 ? 0.0 : capan;
            pp = (double)(third * x1 + geop);
            ratio = x2d / x2dtr;
            for (int a = 0; a < 3; a++) sig[a] = (double)(s_tr[a] * ratio) + pp;
            for (int a = 0; a < 3; a++) sig[3 + a] = (double)(g_sh[a] * ratio);
            break;
        case 6: case 7: {   // :357-398 tension corner / cut-off with the cap moving
            x1 = t;
            x2d = imode == 6 ? fet : x2dtr;
            const float x1old = (float)(sig[0] + sig[1] + sig[2]);
            const float devp = (float)(dd - (double)(x1 - x1old) / (3.0 * k.bulk));
            const float xhat = cap_minf(devpb, devp);
            capa = capa + (double)devp * ((double)capal - capan) / (double)xhat;
            capa = capa > (double)(0.0f * unit) ? capa : (double)(0.0f * unit);
            pp = (double)(third * x1 + geop);
            if (imode == 6) {
                ratio = x2d / x2dtr;
                for (int a = 0; a < 3; a++) sig[a] = (double)(s_tr[a] * ratio) + pp;
                for (int a = 0; a < 3; a++) sig[3 + a] = (double)(g_sh[a] * ratio);
            } else {
                for (int a = 0; a < 3; a++) sig[a] = (double)s_tr[a] + pp;
                for (int a = 0; a < 3; a++) sig[3 + a] = (double)g_sh[a];
            }
            break;
        }
        case 8: case 9:   // :405-442 failure surface, cap not moving / cap corner
            if (imode == 8) {
                x1 = x1tr;
                fe = k.alpha - k.gama * cap_expf(-(k.beta * x1)) + k.theta * x1;
                x2d = cap_minf(fe, tmises);
            } else {
                x1 = elcapn;
                x2d = k.alpha - k.gama * cap_expf(-(k.beta * x1)) + k.theta * x1;
            }
            capa = capan;
            ratio = x2d / x2dtr;
            pp = (double)(third * x1 + geop);
            for (int a = 0; a < 3; a++) sig[a] = (double)(s_tr[a] * ratio) + pp;
            for (int a = 0; a < 3; a++) sig[3 + a] = (double)(g_sh[a] * ratio);
            break;
        case 10: {   // :445-624 failure surface with the cap moving
            for (int a = 0; a < 6; a++) sig[a] = (double)gt[a];
            capa = capan;
            x1 = x1tr;
            x2d = x2dtr;
            float sv[6];
            for (int a = 0; a < 3; a++) sv[a] = (float)(sig[a] - (double)(x1tr / 3.0f));
            for (int a = 3; a < 6; a++) sv[a] = (float)sig[a];
            if (k.ivec == 0) {   // feit: until |f| <= 1e-8, at most 100 iterations
                const float ftol = 1.0e-8f;
                bool conv = false;
                for (int it = 0; it < 100 && !conv; it++) {
                    cap_fail_step(k, capa, sig, x1, x2d, sv);
                    const float f = x2d - (k.alpha - k.gama * cap_expf(-(k.beta * x1)) + k.theta * x1);
                    conv = fabsf(f) <= ftol;
                }
                if (!conv) return -1;
            } else {
                for (int it = 0; it < 4; it++) cap_fail_step(k, capa, sig, x1, x2d, sv);
            }
            if (x1 > elcapn) x1 = elcapn;
            if (ltype == 2 || capan == 0.0) capa = capan;
            if ((double)x1 > capa) capa = (double)x1;
            if (ltype == 1 && capan >= 0.0) capa = capa > (double)(0.0f * unit) ? capa : (double)(0.0f * unit);
            fe = k.alpha - k.gama * cap_expf(-(k.beta * x1)) + k.theta * x1;
            x2d = cap_minf(fe, tmises);
            ratio = x2d / x2dtr;
            pp = (double)(third * x1 + geop);
            for (int a = 0; a < 3; a++) sig[a] = (double)(s_tr[a] * ratio) + pp;
            for (int a = 0; a < 3; a++) sig[3 + a] = (double)(g_sh[a] * ratio);
            break;
        }
        case 11: {   // :626-797 cap surface
            for (int a = 0; a < 6; a++) sig[a] = (double)gt[a];
            capa = capan;
            x1 = x1tr;
            x2d = x2dtr;
            float sv[6], xkc = 0.0f, elcap = 0.0f;
            for (int a = 0; a < 3; a++) sv[a] = (float)(sig[a] - (double)(x1tr / 3.0f));
            for (int a = 3; a < 6; a++) sv[a] = (float)sig[a];
            if (k.ivec == 0) {   // fcit: until |f| <= 1e-8, at most 100 iterations
                const float ftol = 1.0e-8f;
                bool conv = false;
                for (int it = 0; it < 100 && !conv; it++) {
                    cap_cap_step(k, capa, sig, x1, x2d, sv, xkc, elcap);
                    const float f = x2d * x2d + (x1 - elcap) * (x1 - elcap) / (k.r * k.r) -
                                    (xkc - elcap) * (xkc - elcap) / (k.r * k.r);
                    conv = fabsf(f) <= ftol;
                }
                if (!conv) return -1;
            } else {
                for (int it = 0; it < 4; it++) cap_cap_step(k, capa, sig, x1, x2d, sv, xkc, elcap);
            }
            if (x1 < elcap) {
                x1 = elcap + fabsf(x1 - elcap);
                for (int a = 0; a < 3; a++) sig[a] = (double)(sv[a] + x1 / 3.0f);
            }
            break;
        }
        default: break;   // no mode: the stress stays the rotated one
    }
    s.sig[S11] = -sig[0]; s.sig[S22] = -sig[1]; s.sig[S33] = -sig[2];
    s.sig[S12] = -sig[3]; s.sig[S13] = -sig[4]; s.sig[S23] = -sig[5];
    s.rod[S11] = -dd1; s.rod[S22] = -dd2; s.rod[S33] = -dd3;
    s.rod[S12] = -(dd4 * 2.0); s.rod[S13] = -(dd5 * 2.0); s.rod[S23] = -(dd6 * 2.0);
    s.p = -(othird * (s.sig[S11] + s.sig[S22] + s.sig[S33]));
    return imode;
}

// initel, mcm_sets4.f90:43-94 (implicit double precision): the initial cap
// parameter from the initial cap position xint, and the zero of the failure
// surface (the tension cut-off fcut), each by Newton iteration to the
// single-precision tolerance 1.0e-3.  The reference warns past 200
// iterations and goes on; a set-up that needs more than 100000 is refused.
inline int cap_initel(double xint, double alpha, double gama, double beta, double theta, double r, double* capint_out,
                      double* fcut_out) {
    const double tol = (double)1.0e-3f, tiny = (double)1.e-15f;
    double capint = 0.0, delcap = 1.e10, delf = 1.e10, fcut = 0.0;
    for (long it = 0;; it++) {
        if (it > 100000) return -1;
        const double fe = alpha - gama * exp(-(beta * capint)) + theta * capint;
        const double psik = xint - capint - r * fe;
        const double err1 = fabs(psik), err2 = fabs(delcap) / (fabs(capint) + tiny);
        if (!(err1 > tol && err2 > tol)) break;
        const double fep = beta * gama * exp(-(beta * capint)) + theta;
        const double dpsik = -1.0 - r * fep;
        delcap = psik / dpsik;
        capint = capint - delcap;
    }
    for (long it = 0;; it++) {
        if (it > 100000) return -1;
        const double fe = alpha - gama * exp(-(beta * fcut)) + theta * fcut;
        const double err1 = fabs(fe), err2 = fabs(delf) / (fabs(fcut) + tiny);
        if (!(err1 > tol && err2 > tol)) break;
        const double fep = gama * beta * exp(-(beta * fcut)) + theta;
        delf = fe / fep;
        fcut = fcut - delf;
    }
    *capint_out = capint;
    *fcut_out = fcut;
    return 0;
}

// mcm_sets4.f90:19-41: lambda + 2G in cm(5,2) from the bulk and shear
// moduli (only the second tempNi of the reference counts), the initial cap
// parameter in cm(3,3), the failure surface's zero in cm(3,4)
inline int cap_sets4(double* cm) {
#define C(a, b) cm[((a) - 1) + 5 * ((b) - 1)]
    const double tempE = (9.0 * C(1, 1) * C(1, 2)) / (3.0 * C(1, 1) + C(1, 2));
    const double tempNi = (3.0 * C(1, 1) - tempE) / (6.0 * C(1, 1));
    const double q1 = tempE * tempNi / ((1.0 + tempNi) * (1.0 - 2.0 * tempNi));
    const double q2 = tempE * 0.5 / (1.0 + tempNi);
    const double q3 = q1 + 2.0 * q2;
    C(5, 2) = q3;
    double capint, fcut;
    if (cap_initel(C(2, 5), C(1, 3), C(1, 5), C(1, 6), C(1, 4), C(2, 1), &capint, &fcut)) return -1;
    C(3, 3) = capint;
    C(3, 4) = fcut;
#undef C
    return 0;
}
