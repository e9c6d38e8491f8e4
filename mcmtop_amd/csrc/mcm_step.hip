// mcm_step.hip -- pair and per-particle kernels of one MCM step on gfx950.
//
// Two fused gather kernels carry the whole step:
//   k_strain_const : strain rate (mcm_strain / mcm_oldstrain / mcm_corr_strain)
//                    + mcm_update_h + mcm_constitutive (rhoupd, stress rotation,
//                    material model, EOS, nodal AV, energy) + calccritts, with a
//                    last-block reduction that advances dt and ptime on the device
//                    (mcm_solution.f90:58-62);
//   k_momentum     : momentum (mcm_oldacceleration / mcm_corr_acceleration)
//                    + mcm_boundcond + mcm_calctotale + mcm_update_velocity +
//                    mcm_move (+ mcm_check_sym_pen), with a last-block reduction
//                    of the coordinate bounds and energies.
// Both are gather-only: every thread owns one particle and reads its
// neighbours; nothing is scattered, so results do not depend on scheduling.
// Reductions use per-block partials reduced in a fixed order.
//
// The corrected (normalised) kernels walk the reference's three neighbour
// passes in its order (CORR 1).  A variant (CORR 2) assembles B, the
// corrected gradient contraction and the deltasig contraction from
// single-pass moments instead: the correction is linear in the pass-1 sums,
// so it is the same mathematics with different rounding.
#pragma once
#include "mcm_nbr.hip"

// MODE_FUSED: whole stage in one kernel; MODE_STRAIN / MODE_CONST /
// MODE_MOMENTUM: one reference subroutine per launch (granular entry
// points); MODE_FSTRAIN + MODE_FCONST: the fused step's strain gather, which
// stages rod, spin and tracerod, then update_h + constitutive + dt in a
// kernel of its own, so that the constitutive code's register demand does
// not limit the occupancy of the gathers.
enum { MODE_FUSED = 0, MODE_STRAIN = 1, MODE_CONST = 2, MODE_MOMENTUM = 1, MODE_FCONST = 3, MODE_FSTRAIN = 4,
       MODE_FVEL = 5 };   // MODE_FVEL: fused momentum up to the velocity update, the move follows XSPH

// ------------------------------------------------------------- reductions
// Block-wide reduction (OP 0 sum, 1 min, 2 max) in a fixed order: butterfly
// within each wavefront, then the wavefronts in index order.  Every thread
// of the block must call it; the result is valid in every thread.
template <int OP>
__device__ __forceinline__ double block_reduce(double v) {
    __shared__ double sw[MCM_BLOCK / 64];
    v = OP == 0 ? wave_sum(v) : (OP == 1 ? wave_min(v) : wave_max(v));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sw[wid] = v;
    __syncthreads();
    double r = sw[0];
    for (int w = 1; w < MCM_BLOCK / 64; w++) r = OP == 0 ? r + sw[w] : (OP == 1 ? fmin(r, sw[w]) : fmax(r, sw[w]));
    return r;
}

// Final fold of per-block partials red[b*stride + k], b < nblk, by one
// workgroup launched after the kernel that wrote them (the kernel boundary
// makes every block's partials visible; no device-wide fence per block).
// Thread t combines blocks t, t + BLOCK, ... in order, then the block
// reduction combines the threads in a fixed order.  Every thread calls it.
template <int OP>
__device__ __forceinline__ double fold(const double* red, int stride, int k, int nblk, double init) {
    double v = init;
#pragma unroll 4
    for (int b = threadIdx.x; b < nblk; b += MCM_BLOCK) {
        const double x = red[(long)b * stride + k];
        v = OP == 0 ? v + x : (OP == 1 ? fmin(v, x) : fmax(v, x));
    }
    return block_reduce<OP>(v);
}

// ----------------------------------------------------------- state I/O
__device__ __forceinline__ void load_state(const DevState& d, int i, int cur, PState& s) {
    s.mass = d.mass[i];
    s.rho = d.rbuf[cur][i];
    s.h = d.hbuf[cur][i];
    s.hold = d.hold[i];
    s.h0 = d.h0[i];
    s.rho0 = d.rho0[i];
    s.p = d.p[i];
    s.e = d.e[i];
    s.c = d.c[i];
    s.vabs = d.vabs[i];
    s.pcut = d.pcut[i];
    s.p_cut = d.p_cut[i];
    s.efps = d.efps[i];
    s.fail = d.fail[i];
    s.q = d.q[i];
    s.einc = 0.0;
    s.etry = 0.0;
    for (int k = 0; k < 6; k++) s.sig[k] = d.sig[k][i];
}

// -------------------------------------------------------------- strain
// Neighbour records (rs_at, rm_at): the strain passes read rs[u]
// = {x, y, z, vx, vy, vz, m/rho, h}, whose position is the strain pass's
// position of u: x - v dt / 2 for a real particle, the unshifted x of a
// ghost (mcm_strain.f90:27-31; written by the list build), and the momentum
// passes rm[u] = {x, y, z, h, m, m/rho, A11, A22, A33, A12, A13, A23}.  A
// pass loads only the planes it uses (the others are dead loads).
struct SRec { double x, y, z, vx, vy, vz, vol, h; };
// ASYM: runs with a cap-model material, whose stress is not symmetric; the
// record then also holds A21, A31, A32 from rml[u]
template <bool ASYM> struct MRecT { double x, y, z, h, m, vol, A[6]; };
template <> struct MRecT<true> { double x, y, z, h, m, vol, A[6], AL[3]; };
using MRec = MRecT<false>;

__device__ __forceinline__ SRec load_srec(const double* rs, int j) {
    const double2* r = reinterpret_cast<const double2*>(rs) + RS_SLOTS * (long)j;
    const double2 a = r[0], b = r[1], c = r[2], e = r[3];
    return SRec{a.x, a.y, b.x, c.y, e.x, e.y, c.x, b.y};
}
__device__ __forceinline__ MRec load_mrec(const double* rm, int j) {
    const double2* r = reinterpret_cast<const double2*>(rm) + RM_SLOTS * (long)j;
    const double2 a = r[0], b = r[1], c = r[2], e = r[3], f = r[4], h = r[5];
    return MRec{a.x, a.y, b.x, b.y, c.x, c.y, {e.x, e.y, f.x, f.y, h.x, h.y}};
}
__device__ __forceinline__ MRecT<true> load_mreca(const double* rm, const double* rml, int j) {
    const MRec a = load_mrec(rm, j);
    const double2* l = reinterpret_cast<const double2*>(rml) + 2 * (long)j;
    const double2 p = l[0], q = l[1];
    return MRecT<true>{a.x, a.y, a.z, a.h, a.m, a.vol, {a.A[0], a.A[1], a.A[2], a.A[3], a.A[4], a.A[5]},
                       {p.x, p.y, q.x}};
}
__device__ __forceinline__ SRec load_srec(const DevState& d, int j) { return load_srec(d.rs, j); }
__device__ __forceinline__ MRec load_mrec(const DevState& d, int j) { return load_mrec(d.rm, j); }
template <bool ASYM> __device__ __forceinline__ MRecT<ASYM> load_mrec_t(const DevState& d, int j);
template <> __device__ __forceinline__ MRec load_mrec_t<false>(const DevState& d, int j) { return load_mrec(d.rm, j); }
template <> __device__ __forceinline__ MRecT<true> load_mrec_t<true>(const DevState& d, int j) {
    return load_mreca(d.rm, d.rml, j);
}

// records by type; base2 is the second array of a two-array record (rml)
template <typename Rec> __device__ __forceinline__ Rec load_rec(const double* base, const double* base2, int j);
template <> __device__ __forceinline__ SRec load_rec<SRec>(const double* base, const double*, int j) {
    return load_srec(base, j);
}
template <> __device__ __forceinline__ MRec load_rec<MRec>(const double* base, const double*, int j) {
    return load_mrec(base, j);
}
template <> __device__ __forceinline__ MRecT<true> load_rec<MRecT<true>>(const double* base, const double* base2, int j) {
    return load_mreca(base, base2, j);
}

// Walks one particle's list in order with the next entry's record loaded
// one entry ahead and its index two entries ahead, so that the gathers of a
// thread overlap its pair arithmetic.  Past the end it re-reads the
// particle's own record (in bounds, never used).
template <typename Rec>
struct NbrStream {
    const int* lst;
    const double* base;
    const double* base2;
    int n, k, self, jn, jnn;
    Rec rn;
    // list slots are read unconditionally (slots 0..ncap-1 are allocated),
    // values past the end are replaced by self before use
    __device__ __forceinline__ NbrStream(const int* l, const double* b, int n_, int self_, const double* b2 = nullptr)
        : lst(l), base(b), base2(b2), n(n_), k(0), self(self_) {
        const int l0 = lst[0], l1 = lst[n > 1 ? 64 : 0];
        jn = n > 0 ? l0 : self;
        rn = load_rec<Rec>(base, base2, jn);
        jnn = n > 1 ? l1 : self;
    }
    __device__ __forceinline__ void next(int& j, Rec& r) {
        j = jn;
        r = rn;
        jn = jnn;
        rn = load_rec<Rec>(base, base2, jn);
        const int kk = k + 2 < n ? k + 2 : k;
        const int v = lst[(long)kk * 64];
        jnn = k + 2 < n ? v : self;
        k++;
    }
};

// plain strain rate, mcm_oldstrain.f90:34-64, positions half a step back
// for real particles only (mcm_strain.f90:27-31)
template <bool FAST>
__device__ __forceinline__ void strain_plain(const DevState& d, int i, int t, const double xi[3],
                                             const double vi[3], double hi, double g[3][3]) {
    const int n = d.ntot[i];
    const int* lst = nbr_row(d, t);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) g[a][b] = 0.0;
    NbrStream<SRec> st(lst, d.rs, n, i);
#pragma unroll 2
    for (int k = 0; k < n; k++) {
        int j;
        SRec r;
        st.next(j, r);
        double xj0 = r.x, xj1 = r.y, xj2 = r.z;
        double havg = 0.5 * (hi + r.h);
        double dwdx[3];
        pair_grad<FAST>(xi[0] - xj0, xi[1] - xj1, xi[2] - xj2, havg, d.pi, dwdx);
        double dv0 = r.vx - vi[0], dv1 = r.vy - vi[1], dv2 = r.vz - vi[2];
        double volj = r.vol;
        g[0][0] = g[0][0] + volj * dv0 * dwdx[0];
        g[0][1] = g[0][1] + volj * dv0 * dwdx[1];
        g[0][2] = g[0][2] + volj * dv0 * dwdx[2];
        g[1][0] = g[1][0] + volj * dv1 * dwdx[0];
        g[1][1] = g[1][1] + volj * dv1 * dwdx[1];
        g[1][2] = g[1][2] + volj * dv1 * dwdx[2];
        g[2][0] = g[2][0] + volj * dv2 * dwdx[0];
        g[2][1] = g[2][1] + volj * dv2 * dwdx[1];
        g[2][2] = g[2][2] + volj * dv2 * dwdx[2];
    }
}

// kernel value and gradient of one pair of the corrected passes, centred at
// a (xij = a - b), with the neighbour's half-step position
template <bool FAST>
__device__ __forceinline__ void corr_pair(const SRec& r, const double xi[3], double hi, double pi,
                                          double xj[3], double& w, double dwdx[3]) {
    xj[0] = r.x;
    xj[1] = r.y;
    xj[2] = r.z;
    double havg = 0.5 * (hi + r.h);
    pair_wgrad<FAST>(xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2], havg, pi, w, dwdx);
}

// (grad W sw - W sg) / sw^2 of one pair: unchecked quotients when the
// particle's sums allow them on the fast path (fast_sums), else div3
__device__ __forceinline__ void corr_grad(const Divisor& isw2, bool unchecked, double sw, const double sg[3],
                                          double w, const double dwdx[3], double cg[3]) {
    const double n0 = dwdx[0] * sw - w * sg[0], n1 = dwdx[1] * sw - w * sg[1], n2 = dwdx[2] * sw - w * sg[2];
    if (unchecked) {
        cg[0] = isw2.quot(n0);
        cg[1] = isw2.quot(n1);
        cg[2] = isw2.quot(n2);
    } else {
        isw2.div3(n0, n1, n2, cg);
    }
}

// corrected strain rate, mcm_corr_strain.f90:37-146, in the reference's three
// passes and evaluation order
template <bool FAST>
__device__ __forceinline__ void strain_corr(const DevState& d, int i, int t, const double xi[3], double hi,
                                            double f[3][3]) {
    const int n = d.ntot[i];
    const int* lst = nbr_row(d, t);
    double sw = 0.0, sg[3] = {0.0, 0.0, 0.0};
    NbrStream<SRec> s1(lst, d.rs, n, i);
#pragma unroll 2
    for (int k = 0; k < n; k++) {
        int j;
        SRec r;
        s1.next(j, r);
        double xj[3], w, dwdx[3];
        corr_pair<FAST>(r, xi, hi, d.pi, xj, w, dwdx);
        sw = sw + r.vol * w;
        for (int l = 0; l < 3; l++) sg[l] = sg[l] + r.vol * dwdx[l];
    }
    const double sw2 = sw * sw;
    const Divisor isw2(sw2);
    const bool unchecked = FAST && __all(fast_sums(sw, sg));
    double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, invB[9];
    NbrStream<SRec> s2(lst, d.rs, n, i);
#pragma unroll 2
    for (int k = 0; k < n; k++) {
        int j;
        SRec r;
        s2.next(j, r);
        double xj[3], w, dwdx[3], cg[3];
        corr_pair<FAST>(r, xi, hi, d.pi, xj, w, dwdx);
        corr_grad(isw2, unchecked, sw, sg, w, dwdx, cg);
        for (int m = 0; m < 3; m++)
            for (int nn = 0; nn < 3; nn++) B[nn + 3 * m] = B[nn + 3 * m] + r.vol * (xj[nn] - xi[nn]) * cg[m];
    }
    inverse3(B, invB);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) f[a][b] = 0.0;
    NbrStream<SRec> s3(lst, d.rs, n, i);
    for (int k = 0; k < n; k++) {
        int j;
        SRec r;
        s3.next(j, r);
        double xj[3], w, dwdx[3], cg[3], g[3];
        corr_pair<FAST>(r, xi, hi, d.pi, xj, w, dwdx);
        corr_grad(isw2, unchecked, sw, sg, w, dwdx, cg);
        for (int q = 0; q < 3; q++) g[q] = invB[q] * cg[0] + invB[q + 3] * cg[1] + invB[q + 6] * cg[2];
        double vj[3] = {r.vx, r.vy, r.vz};
        for (int nn = 0; nn < 3; nn++)
            for (int m = 0; m < 3; m++) f[nn][m] = f[nn][m] + r.vol * vj[nn] * g[m];
    }
}

// loop unrolling of the one-pass pair loops (strain, momentum)
#ifndef MCM_UNROLL_S1
#define MCM_UNROLL_S1 1
#endif
#ifndef MCM_UNROLL_M1
#define MCM_UNROLL_M1 2
#endif

// W and grad W of one pair of the one-pass kernels: pair_wgrad_1p on the
// fast path, else the checked functions
template <bool FAST>
__device__ __forceinline__ void pair_wgrad_one_pass(double xij0, double xij1, double xij2, double havg, double pi,
                                                    double ipi, double& w, double dwdx[3]) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (FAST) {
        pair_wgrad_1p(xij0, xij1, xij2, havg, ipi, w, dwdx);
        return;
    }
#endif
    (void)ipi;
    pair_wgrad<false>(xij0, xij1, xij2, havg, pi, w, dwdx);
}

// corrected strain rate in one pass: the correction is linear in the pass-1
// sums, so B and the contracted gradient follow from single-pass moments
// (same result up to rounding; the fast variant)
template <bool FAST>
__device__ __forceinline__ void strain_corr_fast(const DevState& d, int i, int t, const double xi[3],
                                                 double hi, double f[3][3]) {
#pragma clang fp contract(fast)
    const int n = d.ntot[i];
    const double ipi = 1.0 / d.pi;
    const int* lst = nbr_row(d, t);
    double sw = 0.0, sg[3] = {0, 0, 0}, P[3][3] = {{0}}, Pw[3] = {0, 0, 0}, Q[3][3] = {{0}}, Qw[3] = {0, 0, 0};
    NbrStream<SRec> st(lst, d.rs, n, i);
#pragma unroll MCM_UNROLL_S1
    for (int k = 0; k < n; k++) {
        int j;
        SRec r;
        st.next(j, r);
        double xj[3] = {r.x, r.y, r.z}, w, dwdx[3];
        pair_wgrad_one_pass<FAST>(xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2], 0.5 * (hi + r.h), d.pi, ipi, w,
                                  dwdx);
        double vw = r.vol * w;
        double vj[3] = {r.vx, r.vy, r.vz};
        sw += vw;
        double vg[3] = {r.vol * dwdx[0], r.vol * dwdx[1], r.vol * dwdx[2]};
        for (int a = 0; a < 3; a++) {
            sg[a] += vg[a];
            double dxa = xj[a] - xi[a];
            Pw[a] += dxa * vw;
            Qw[a] += vj[a] * vw;
            for (int b = 0; b < 3; b++) {
                P[a][b] += dxa * vg[b];
                Q[a][b] += vj[a] * vg[b];
            }
        }
    }
    double sw2 = sw * sw;
    double B[9], invB[9];
    for (int m = 0; m < 3; m++)
        for (int a = 0; a < 3; a++) B[a + 3 * m] = (P[a][m] * sw - Pw[a] * sg[m]) / sw2;
    inverse3(B, invB);
    for (int a = 0; a < 3; a++) {
        double t[3];
        for (int q = 0; q < 3; q++) t[q] = (Q[a][q] * sw - Qw[a] * sg[q]) / sw2;
        for (int m = 0; m < 3; m++) f[a][m] = invB[m] * t[0] + invB[m + 3] * t[1] + invB[m + 6] * t[2];
    }
}

// ------------------------------------------------------- k_strain_const
// After the strain pass of particle i: the reference moves x back by the
// same half step, (x - t) + t (mcm_strain.f90:51-58); x1 and the momentum
// record's position are written.  Returns x1z.
__device__ __forceinline__ double strain_finish(const DevState& d, int i, double dt, const double xi[3],
                                                const double vi[3]) {
    double x1[3];
    for (int a = 0; a < 3; a++) {
        x1[a] = xi[a] + 0.5 * vi[a] * dt;
        d.x1[a][i] = x1[a];
    }
    *rm_at(d, 0, i) = make_double2(x1[0], x1[1]);
    if (!(fast_coord(x1[0]) & fast_coord(x1[1]) & fast_coord(x1[2]))) atomicOr(d.range, (int)RANGE_RM);
    return x1[2];
}

// mcm_update_h + mcm_constitutive for particle i with the strain rate in s
// (rod, spin, tracerod); MODE_CONST reads the state staged by the strain and
// update_h stages.  Returns the particle's critical timestep, the updated h
// (0 when not updated) and sets eos_err.
template <int MODE, bool DIAG, int MODELS = 0>
__device__ __forceinline__ void constitutive_particle(const DevState& d, int i, int cur, double dt, PState& s,
                                                      double x1z, double& crit, double& hm, int& eos_err) {
    const int nxt = cur ^ 1;
    load_state(d, i, MODE == MODE_CONST ? nxt : cur, s);
    const int mi = d.pmat[i];
    const MatParams& m = d.mat[mi];
    if ((MODE == MODE_FUSED || MODE == MODE_FCONST) && d.h_opt > 0) {
        if (update_h(s, d.h_opt, dt)) hm = s.h;
    }
    rhoupd(s, m, dt);
    constexpr bool has3 = MODELS == 0 || (MODELS & (1 << 3)), has9 = MODELS == 0 || (MODELS & (1 << 9));
    if (has3 && m.model == 3) for (int k = 0; k < 6; k++) s.alfa[k] = d.alfa[k][i];
    if (has9 && m.model == 9) for (int k = 0; k < 6; k++) s.s[k] = d.sd[k][i];
    double defps_prev = d.defps_prev[i];
    int active = 0;
    // lower triangle of sigma: 21, 31, 32
    double sl[3];
    if ((MODELS == 0 || (MODELS & (1 << 4))) && m.model == 4) {
        // mcm_constitutive.f90:49-54: stress_rot1, f3dm25, aviscosity, lieupd.
        // The rotation reads sigma 12, 23 and 31 (mcm_stress_rot1.f90:33-41,
        // 163-171) and writes a symmetric tensor; the cap model then writes
        // the upper triangle only
        double capa = d.capa[i];
        s.sig[S13] = d.sigl[1][i];
        stress_rot1(s, dt);
        sl[0] = s.sig[S12]; sl[1] = s.sig[S13]; sl[2] = s.sig[S23];
        int mode = f3dm25(s, cap_const(m), dt, capa);
        if (mode < 0) {
            eos_err |= 2;
            mode = 0;
        }
        nodalav(s, m);
        lieupd(s, dt);
        d.capa[i] = capa;
        d.ptype[i] = mode;
    } else {
        active = constitutive_one<MODELS>(s, m, mi + 1, dt, defps_prev, &eos_err);
        sl[0] = s.sig[S12]; sl[1] = s.sig[S13]; sl[2] = s.sig[S23];
    }
    s.critts = critts(s, m, d.tcrit_opt, d.mindist[i]);
    crit = s.critts;
    // write back
    d.rbuf[nxt][i] = s.rho;
    d.hbuf[nxt][i] = s.h;
    if (d.h_opt > 0) d.hold[i] = s.hold;
    for (int k = 0; k < 6; k++) d.sig[k][i] = s.sig[k];
    if (has3 && m.model == 3) for (int k = 0; k < 6; k++) d.alfa[k][i] = s.alfa[k];
    if ((has9 && m.model == 9) || (DIAG && m.model == 10)) for (int k = 0; k < 6; k++) d.sd[k][i] = s.s[k];
    if (DIAG && has3 && m.model == 3) for (int k = 0; k < 6; k++) d.sd[k][i] = s.s[k];
    d.p[i] = s.p;
    d.e[i] = s.e;
    d.q[i] = s.q;
    d.efps[i] = s.efps;
    if (m.model == 9 || m.model == 10) {
        d.c[i] = s.c;
        d.pcut[i] = s.pcut;
        d.p_cut[i] = s.p_cut;
        d.fail[i] = s.fail;
        d.defps_prev[i] = defps_prev;
    }
    // momentum record: x after the strain pass, new h, mass, volume and
    // (sigma - q)/rho^2 (mcm_oldacceleration.f90:79-80, per particle)
    double r2 = s.rho * s.rho;
    *rm_at(d, 1, i) = make_double2(x1z, s.h);
    if (!fast_h(s.h)) atomicOr(d.range, (int)RANGE_RM);
    *rm_at(d, 2, i) = make_double2(s.mass, s.mass / s.rho);
    *rm_at(d, 3, i) = make_double2((s.sig[S11] - s.q) / r2, (s.sig[S22] - s.q) / r2);
    *rm_at(d, 4, i) = make_double2((s.sig[S33] - s.q) / r2, (s.sig[S12] - 0.0) / r2);
    *rm_at(d, 5, i) = make_double2((s.sig[S13] - 0.0) / r2, (s.sig[S23] - 0.0) / r2);
    if (d.asym) {
        for (int k = 0; k < 3; k++) d.sigl[k][i] = sl[k];
        double2* l = reinterpret_cast<double2*>(d.rml) + 2 * (long)i;
        l[0] = make_double2((sl[0] - 0.0) / r2, (sl[1] - 0.0) / r2);
        l[1] = make_double2((sl[2] - 0.0) / r2, 0.0);
    }
    if (DIAG) {
        for (int k = 0; k < 6; k++) d.rod[k][i] = s.rod[k];
        for (int k = 0; k < 3; k++) d.spin[k][i] = s.spin[k];
        d.tracerod[i] = s.tracerod;
        d.einc[i] = s.einc;
        d.etry[i] = s.etry;
        d.rhoold[i] = s.rhoold;
        d.critts_p[i] = s.critts;
        d.yield[i] = active;
    }
}

// energy and coordinate-bound partials: [cmin 3][cmax 3][ke nm][ie nm][mom 3nm]
#define RED_C (6 + 5 * MCM_MAX_MAT)

// partial k of block lb: the partials are stored value by value, [k][nblk],
// so that the one-workgroup folds read them coalesced
__device__ __forceinline__ double* red_at(const DevState& d, int lb, int k) { return d.red + (long)k * d.nblk + lb; }

// One thread's share of the step's end-of-step reductions: its particle's
// new position (coordinate bounds) and its terms of mcm_calctotale.f90:33-51
// for its material mi (mi < 0: no particle).
struct CPart {
    double x[3];
    double ke, ie, mom[3];
    int mi;
};

__device__ __forceinline__ void init_c_part(CPart& c) {
    for (int a = 0; a < 3; a++) { c.x[a] = 0.0; c.mom[a] = 0.0; }
    c.ke = 0.0; c.ie = 0.0; c.mi = -1;
}

__device__ __forceinline__ void energy_terms(CPart& c, int mi, double mass, double vabs, double e, const double v[3]) {
    c.mi = mi;
    c.ke = 0.5 * mass * (vabs * vabs);
    c.ie = e;
    for (int a = 0; a < 3; a++) c.mom[a] = mass * fabs(v[a]);
}

// Block partials of every block (folded by k_c_fold): the coordinate
// bounds [cmin 3][cmax 3] (sentinels without bounds) ...
__device__ void reduce_bound_partials(const DevState& d, const CPart& c, bool with_bounds, int lb) {
    const bool has = c.mi >= 0;
    for (int a = 0; a < 3; a++) {
        double lo = block_reduce<1>(has && with_bounds ? c.x[a] : 1.0e20);
        double hi = block_reduce<2>(has && with_bounds ? c.x[a] : -1.0e20);
        if (threadIdx.x == 0) { *red_at(d, lb, a) = lo; *red_at(d, lb, 3 + a) = hi; }
    }
}

// ... and the energies [ke nm][ie nm][mom 3nm].  The fused step takes them
// in the constitutive pass, in particle order (their terms use e after the
// update and v, |v| before the momentum pass), so that the sums do not
// depend on the momentum pass's particle order.  Every thread of the block
// calls it.
__device__ void reduce_energy_partials(const DevState& d, const CPart& c, int lb) {
    const int nm = d.nummat;
    for (int m = 0; m < nm; m++) {
        const bool mine = c.mi == m;
        double ke = block_reduce<0>(mine ? c.ke : 0.0);
        double ie = block_reduce<0>(mine ? c.ie : 0.0);
        if (threadIdx.x == 0) { *red_at(d, lb, 6 + m) = ke; *red_at(d, lb, 6 + nm + m) = ie; }
        for (int a = 0; a < 3; a++) {
            double mo = block_reduce<0>(mine ? c.mom[a] : 0.0);
            if (threadIdx.x == 0) *red_at(d, lb, 6 + 2 * nm + 3 * m + a) = mo;
        }
    }
}

__device__ void reduce_c_partials(const DevState& d, const CPart& c, bool with_bounds, int lb) {
    reduce_bound_partials(d, c, with_bounds, lb);
    reduce_energy_partials(d, c, lb);
}

// Block min of critts and max of the updated h into the first two slots of
// the block's partials (folded by k_strain_fold, before the momentum pass
// writes its coordinate bounds there).  Every thread of the block calls it.
__device__ __forceinline__ void strain_block_partials(const DevState& d, int lb, double crit, double hm, int eos_err) {
    if (eos_err & 1) atomicOr(&d.sc->error_flags, 16);
    if (eos_err & 2) {   // the cap model's iterated return did not converge (the reference stops)
        atomicOr(&d.sc->error_flags, 128);
        atomicExch(&d.sc->abort, 1);
    }
    double cb = block_reduce<1>(crit);
    double hb = block_reduce<2>(hm);
    if (threadIdx.x == 0) {
        *red_at(d, lb, 0) = cb;
        *red_at(d, lb, 1) = hb;
    }
}

// critts and hmax from the partials of nblk blocks; in the fused step dt and
// ptime advance (mcm_solution.f90:58-62).  One workgroup.
template <int MODE>
__global__ __launch_bounds__(MCM_BLOCK) void k_strain_fold(DevState d_) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    const int nblk = d.nblk;
    Scalars* sc = d.sc;
    double c1 = fold<1>(d.red, 1, 0, nblk, 1.0e20);
    double h1 = fold<2>(d.red + nblk, 1, 0, nblk, 0.0);
    if (threadIdx.x == 0 && !sc->abort) {
        sc->critts = c1;
        if ((MODE == MODE_FUSED || MODE == MODE_FCONST) && d.h_opt > 0) sc->hmax = h1;
        if (MODE == MODE_FUSED || MODE == MODE_FCONST) {
            sc->dtold = sc->dt;
            sc->dt = d.tssf * sc->critts;
            sc->dt = fmin(sc->dt, 1.1 * sc->dtold);
            sc->ptime = sc->ptime + sc->dt;
        }
    }
}

// Thread-per-particle strain (+ constitutive) kernel.  CORR: 0 plain SPH,
// 1 corrected (reference order), 2 corrected (one pass).
template <int MODE, int CORR, bool DIAG, int MODELS = 0>
// Waves per SIMD of the pair kernels (A/B-timed on the 2.1 M-particle bar).
// The gathers are bound by the vector L1 path, and fewer resident waves
// evict fewer of each other's neighbour lines: the one-pass corrected
// kernels (28 and 40 moment sums) run fastest at 2 and 1 waves, the
// three-pass ones at 3 and 2.
#ifndef MCM_WPE_STRAIN1
#define MCM_WPE_STRAIN1 2
#endif
#ifndef MCM_WPE_MOM1
#define MCM_WPE_MOM1 1
#endif
#ifndef MCM_WPE_STRAIN3
#define MCM_WPE_STRAIN3 3
#endif
#ifndef MCM_WPE_MOM3
#define MCM_WPE_MOM3 2
#endif
__global__ __launch_bounds__(MCM_BLOCK) __attribute__((amdgpu_waves_per_eu(
    MODE == MODE_FSTRAIN ? (CORR == 2 ? MCM_WPE_STRAIN1 : MCM_WPE_STRAIN3) : 3))) void k_strain_const(DevState d_,
                                                                                                       int cur) {
    Blk bk;
    const DevState& d = enter<true>(d_, bk);
    const int lb = bk.b;
    const int nxt = cur ^ 1;
    Scalars* sc = d.sc;
    // the strain pass takes the particles in the search's sorted order (the
    // order of the lists, nbr_row); the constitutive pass in id order
    constexpr bool pairs = MODE == MODE_FUSED || MODE == MODE_STRAIN || MODE == MODE_FSTRAIN;
    const int t = lb * MCM_BLOCK + threadIdx.x;
    const bool live = t < d.np && !sc->abort;
    const int i = pairs && live ? d.sorted[t] : t;
    const double dt = sc->dt;
    double crit = 1.0e20, hm = 0.0;
    int eos_err = 0;
    CPart ce;   // energy terms (MODE_FCONST)
    init_c_part(ce);
    if (live) {
        PState s;
        double x1z = 0.0;
        if (MODE == MODE_FUSED || MODE == MODE_STRAIN || MODE == MODE_FSTRAIN) {
            double xr[3], vi[3], xi[3];
            for (int a = 0; a < 3; a++) {
                xr[a] = d.x0[a][i];
                vi[a] = d.v[a][i];
                xi[a] = xr[a] - 0.5 * vi[a] * dt;
            }
            double hi = d.hbuf[cur][i];
            double g[3][3];
            // records within the fast-path ranges (flag written by the search)
            const bool fast = (d.range[0] & RANGE_RS) == 0;
            if (CORR == 1) {
                if (fast) strain_corr<true>(d, i, t, xi, hi, g);
                else strain_corr<false>(d, i, t, xi, hi, g);
            } else if (CORR == 2) {
                if (fast) strain_corr_fast<true>(d, i, t, xi, hi, g);
                else strain_corr_fast<false>(d, i, t, xi, hi, g);
            } else {
                if (fast) strain_plain<true>(d, i, t, xi, vi, hi, g);
                else strain_plain<false>(d, i, t, xi, vi, hi, g);
            }
            rod_spin(g, s);
            x1z = strain_finish(d, i, dt, xi, vi);
            if (MODE == MODE_STRAIN) {
                *rm_at(d, 1, i) = make_double2(x1z, hi);
                if (!fast_h(hi)) atomicOr(d.range, (int)RANGE_RM);
                for (int k = 0; k < 6; k++) d.rod[k][i] = s.rod[k];
                for (int k = 0; k < 3; k++) d.spin[k][i] = s.spin[k];
                d.tracerod[i] = s.tracerod;
                d.hbuf[nxt][i] = hi;
                d.rbuf[nxt][i] = d.rbuf[cur][i];
            }
            if (MODE == MODE_FSTRAIN) {
                for (int k = 0; k < 6; k++) d.rod[k][i] = s.rod[k];
                for (int k = 0; k < 3; k++) d.spin[k][i] = s.spin[k];
                d.tracerod[i] = s.tracerod;
            }
        } else {
            for (int k = 0; k < 6; k++) s.rod[k] = d.rod[k][i];
            for (int k = 0; k < 3; k++) s.spin[k] = d.spin[k][i];
            s.tracerod = d.tracerod[i];
            x1z = d.x1[2][i];
        }
        if (MODE != MODE_STRAIN && MODE != MODE_FSTRAIN)
            constitutive_particle<MODE, DIAG, MODELS>(d, i, cur, dt, s, x1z, crit, hm, eos_err);
        if (MODE == MODE_FCONST) {   // mcm_calctotale.f90:33-51 of the fused step
            const double v[3] = {d.v[0][i], d.v[1][i], d.v[2][i]};
            energy_terms(ce, d.pmat[i], s.mass, s.vabs, s.e, v);
        }
    }
    if (MODE == MODE_STRAIN || MODE == MODE_FSTRAIN) return;
    strain_block_partials(d, lb, crit, hm, eos_err);
    if (MODE == MODE_FCONST) reduce_energy_partials(d, ce, lb);
}

// mcm_update_h.f90:25-54 as a stage of its own (granular path)
__global__ __launch_bounds__(MCM_BLOCK) void k_update_h(DevState d, int cur) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    const int nxt = cur ^ 1;
    Scalars* sc = d.sc;
    double hm = 0.0;
    if (i < d.np && !sc->abort) {
        PState s;
        s.h = d.hbuf[nxt][i]; s.hold = d.hold[i]; s.h0 = d.h0[i]; s.rho0 = d.rho0[i];
        s.rho = d.rbuf[nxt][i]; s.p = d.p[i]; s.pcut = d.pcut[i]; s.tracerod = d.tracerod[i];
        if (update_h(s, d.h_opt, sc->dt)) hm = s.h;
        d.hbuf[nxt][i] = s.h;
        d.hold[i] = s.hold;
    }
    double hb = block_reduce<2>(hm);
    if (threadIdx.x == 0) d.red[blockIdx.x] = hb;
}

__global__ __launch_bounds__(MCM_BLOCK) void k_update_h_fold(DevState d, int nblk) {
    double h1 = fold<2>(d.red, 1, 0, nblk, 0.0);
    if (threadIdx.x == 0) d.sc->hmax = h1;
}

// dt/ptime update of mcm_solution.f90:58-62 (granular path)
__global__ void k_dt_update(DevState d) {
    Scalars* sc = d.sc;
    if (sc->abort) return;
    sc->dtold = sc->dt;
    sc->dt = d.tssf * sc->critts;
    sc->dt = fmin(sc->dt, 1.1 * sc->dtold);
    sc->ptime = sc->ptime + sc->dt;
}

// mcm_update_ghost_stress.f90:28-41: a ghost carries its parent's stress and
// viscosity with every off-diagonal component negated, divided by the
// ghost's own (start-of-step) density
__global__ __launch_bounds__(MCM_BLOCK) void k_ghost_stress(DevState d_, int cur) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    const int nxt = cur ^ 1;
    if (d.sc->abort) return;
    const int ngp = d.sc->ngp;
    for (int g = bk.b * MCM_BLOCK + threadIdx.x; g < ngp; g += bk.n * MCM_BLOCK) {
        const int u = d.np + g, p = d.gpar[g];
        double rg = d.rbuf[nxt][u];
        double r2 = rg * rg;
        double q = d.q[p];
        *rm_at(d, 3, u) = make_double2((d.sig[S11][p] - q) / r2, (d.sig[S22][p] - q) / r2);
        *rm_at(d, 4, u) = make_double2((d.sig[S33][p] - q) / r2, -d.sig[S12][p] / r2);
        *rm_at(d, 5, u) = make_double2(-d.sig[S13][p] / r2, -d.sig[S23][p] / r2);
        if (d.asym) {   // the lower triangle, mirrored the same way
            double2* l = reinterpret_cast<double2*>(d.rml) + 2 * (long)u;
            l[0] = make_double2(-d.sigl[0][p] / r2, -d.sigl[1][p] / r2);
            l[1] = make_double2(-d.sigl[2][p] / r2, 0.0);
        }
    }
}

// ------------------------------------------------------------- momentum
__device__ __forceinline__ int sym_index(int r, int c) {
    return r == c ? r : ((r + c == 1) ? S12 : (r + c == 2 ? S13 : S23));
}

// (sigma - q)/rho^2 of a record at (row, col); below the diagonal the
// lower triangle of a record that has one (21, 31, 32 in AL)
template <bool ASYM>
__device__ __forceinline__ double acomp(const MRecT<ASYM>& r, int row, int col) {
    if constexpr (ASYM) {
        if (row > col) return r.AL[row + col - 1];
    }
    return r.A[sym_index(row, col)];
}
// components held by the corrected sums: the 6 of a symmetric tensor, or
// all 9 (row * 3 + col)
template <bool ASYM> constexpr int kNComp = ASYM ? 9 : 6;
template <bool ASYM>
__device__ __forceinline__ double ccomp(const MRecT<ASYM>& r, int c) {
    if constexpr (ASYM) return acomp<true>(r, c / 3, c % 3);
    else return r.A[c];
}
template <bool ASYM>
__device__ __forceinline__ int cindex(int row, int col) {
    if constexpr (ASYM) return row * 3 + col;
    else return sym_index(row, col);
}

// mcm_oldacceleration.f90:60-89
// a starts at 0, plus the contact repulsion when penalty contact is on
// (mcm_oldacceleration.f90:43-56, mcm_corr_acceleration.f90:38-49)
__device__ __forceinline__ double accel_start(const DevState& d, int i, int k) {
    return d.contacttype > 0 ? 0.0 + d.rep[k][i] : 0.0;
}

template <bool FAST, bool ASYM>
__device__ __forceinline__ void accel_plain(const DevState& d, int i, int t, const MRecT<ASYM>& me, double acc[3]) {
    const int n = d.ntot[i];
    const int* lst = nbr_row(d, t);
    acc[0] = accel_start(d, i, 0); acc[1] = accel_start(d, i, 1); acc[2] = accel_start(d, i, 2);
    NbrStream<MRecT<ASYM>> st(lst, d.rm, n, i, d.rml);
#pragma unroll 2
    for (int k = 0; k < n; k++) {
        int j;
        MRecT<ASYM> r;
        st.next(j, r);
        double havg = 0.5 * (me.h + r.h);
        double dwdx[3];
        pair_grad<FAST>(me.x - r.x, me.y - r.y, me.z - r.z, havg, d.pi, dwdx);
        for (int nn = 0; nn < 3; nn++) {
            double ds = 0.0;
            for (int mm = 0; mm < 3; mm++) ds = ds + (acomp(r, mm, nn) + acomp(me, mm, nn)) * dwdx[mm];
            acc[nn] = acc[nn] + r.m * ds;
        }
    }
}

// kernel value and gradient centred at the neighbour: W(x_j, x_i),
// grad_j W (mcm_corr_acceleration.f90:73-74)
template <bool FAST, class R>
__device__ __forceinline__ void corr_pair_j(const R& r, const R& me, double pi, double& w, double dwdx[3]) {
    double havg = 0.5 * (me.h + r.h);
    pair_wgrad<FAST>(r.x - me.x, r.y - me.y, r.z - me.z, havg, pi, w, dwdx);
}

// mcm_corr_acceleration.f90:37-163 in the reference's three passes: the self
// slot (k == nnbr) contributes nothing to deltasig, the mass of the last
// neighbour multiplies the sum and B^-1 enters transposed
template <bool FAST, bool ASYM>
__device__ __forceinline__ void accel_corr(const DevState& d, int i, int t, const MRecT<ASYM>& me, double acc[3]) {
    const int n = d.ntot[i];
    const int self_k = d.nreal[i] - 1;
    const int* lst = nbr_row(d, t);
    double sw = 0.0, sg[3] = {0.0, 0.0, 0.0}, massj = 0.0;
    int j;
    NbrStream<MRecT<ASYM>> s1(lst, d.rm, n, i, d.rml);
#pragma unroll 2
    for (int k = 0; k < n; k++) {
        MRecT<ASYM> r;
        s1.next(j, r);
        double w, dwdx[3];
        corr_pair_j<FAST>(r, me, d.pi, w, dwdx);
        sw = sw + r.vol * w;
        for (int l = 0; l < 3; l++) sg[l] = sg[l] + r.vol * dwdx[l];
        massj = r.m;
    }
    const double sw2 = sw * sw;
    const Divisor isw2(sw2);
    const bool unchecked = FAST && __all(fast_sums(sw, sg));
    double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, invB[9];
    NbrStream<MRecT<ASYM>> s2(lst, d.rm, n, i, d.rml);
#pragma unroll 2
    for (int k = 0; k < n; k++) {
        MRecT<ASYM> r;
        s2.next(j, r);
        double w, dwdx[3], cg[3];
        corr_pair_j<FAST>(r, me, d.pi, w, dwdx);
        corr_grad(isw2, unchecked, sw, sg, w, dwdx, cg);
        double dx[3] = {r.x - me.x, r.y - me.y, r.z - me.z};
        for (int m = 0; m < 3; m++)
            for (int nn = 0; nn < 3; nn++) B[nn + 3 * m] = B[nn + 3 * m] + r.vol * dx[nn] * cg[m];
    }
    inverse3(B, invB);
    // deltasig(m,n,l) is symmetric in (m,n) unless the stress is not: 6 or
    // 9 components x 3 accumulators
    double D[kNComp<ASYM>][3] = {{0.0}};
    NbrStream<MRecT<ASYM>> s3(lst, d.rm, n, i, d.rml);
#pragma unroll 2
    for (int k = 0; k < n; k++) {
        MRecT<ASYM> r;
        s3.next(j, r);
        if (k == self_k) continue;   // dsig = 0 there
        double w, dwdx[3], cg[3];
        corr_pair_j<FAST>(r, me, d.pi, w, dwdx);
        corr_grad(isw2, unchecked, sw, sg, w, dwdx, cg);
        for (int c = 0; c < kNComp<ASYM>; c++) {
            double dsig = ccomp(r, c) + ccomp(me, c);
            D[c][0] = D[c][0] + dsig * cg[0];
            D[c][1] = D[c][1] + dsig * cg[1];
            D[c][2] = D[c][2] + dsig * cg[2];
        }
    }
    for (int nn = 0; nn < 3; nn++) {
        double a = accel_start(d, i, nn);
        for (int mm = 0; mm < 3; mm++) {
            int c = cindex<ASYM>(nn, mm);
            for (int l = 0; l < 3; l++) a = a + massj * D[c][l] * invB[l + 3 * mm];
        }
        acc[nn] = a;
    }
}

// corrected acceleration in one pass from single-pass moments (fast variant)
template <bool FAST, bool ASYM>
__device__ __forceinline__ void accel_corr_fast(const DevState& d, int i, int t, const MRecT<ASYM>& me, double acc[3]) {
#pragma clang fp contract(fast)
    const int n = d.ntot[i];
    const double ipi = 1.0 / d.pi;
    const int self_k = d.nreal[i] - 1;
    const int* lst = nbr_row(d, t);
    double sw = 0.0, sg[3] = {0, 0, 0}, P[3][3] = {{0}}, Pw[3] = {0, 0, 0};
    double E1[kNComp<ASYM>][3] = {{0}}, E2[kNComp<ASYM>] = {0};
    double mlast = 0.0;
    int j;
    NbrStream<MRecT<ASYM>> st(lst, d.rm, n, i, d.rml);
#pragma unroll MCM_UNROLL_M1
    for (int k = 0; k < n; k++) {
        MRecT<ASYM> r;
        st.next(j, r);
        double w, dwdx[3];
        double dx[3] = {r.x - me.x, r.y - me.y, r.z - me.z};
        pair_wgrad_one_pass<FAST>(dx[0], dx[1], dx[2], 0.5 * (me.h + r.h), d.pi, ipi, w, dwdx);
        double vw = r.vol * w;
        sw += vw;
        for (int a = 0; a < 3; a++) {
            sg[a] += r.vol * dwdx[a];
            Pw[a] += dx[a] * vw;
            for (int b = 0; b < 3; b++) P[a][b] += dx[a] * (r.vol * dwdx[b]);
        }
        if (k != self_k) {
            for (int c = 0; c < kNComp<ASYM>; c++) {
                double ds = ccomp(r, c) + ccomp(me, c);
                E2[c] += ds * w;
                E1[c][0] += ds * dwdx[0];
                E1[c][1] += ds * dwdx[1];
                E1[c][2] += ds * dwdx[2];
            }
        }
        mlast = r.m;
    }
    double sw2 = sw * sw;
    double B[9], invB[9];
    for (int m = 0; m < 3; m++)
        for (int a = 0; a < 3; a++) B[a + 3 * m] = (P[a][m] * sw - Pw[a] * sg[m]) / sw2;
    inverse3(B, invB);
    for (int nn = 0; nn < 3; nn++) {
        double sacc = 0.0;
        for (int mm = 0; mm < 3; mm++) {
            int c = cindex<ASYM>(nn, mm);
            for (int l = 0; l < 3; l++) sacc += (E1[c][l] * sw - E2[c] * sg[l]) / sw2 * invB[l + 3 * mm];
        }
        acc[nn] = accel_start(d, i, nn) + mlast * sacc;
    }
}

__device__ __forceinline__ void boundcond(int bc, double acc[3]) {
    // mcm_boundcond.f90:33-62
    switch (bc) {
        case 1: acc[0] = 0.0; break;
        case 2: acc[1] = 0.0; break;
        case 3: acc[2] = 0.0; break;
        case 4: acc[0] = 0.0; acc[1] = 0.0; break;
        case 5: acc[1] = 0.0; acc[2] = 0.0; break;
        case 6: acc[0] = 0.0; acc[2] = 0.0; break;
        case 7: acc[0] = 0.0; acc[1] = 0.0; acc[2] = 0.0; break;
        default: break;
    }
}

// mcm_check_sym_pen.f90:27-179, per particle
__device__ __forceinline__ void sym_pen(const DevState& d, double x[3], double v[3]) {
    for (int k = 0; k < 3; k++) {
        double b0 = d.bx[k][0], b1 = d.bx[k][1];
        if (d.bcode[k][0] == 1 && x[k] < b0) { x[k] = 2.0 * b0 - x[k]; v[k] = -v[k]; }
        if (d.bcode[k][1] == 1 && x[k] > b1) { x[k] = 2.0 * b1 - x[k]; v[k] = -v[k]; }
        if (d.bcode[k][0] == 2) {
            double span = b1 - b0;
            if (x[k] < b0) x[k] = x[k] + span;
            if (x[k] > b1) x[k] = x[k] - span;
        }
    }
}

// Coordinate bounds (with_bounds: the step's move), per-material energies
// and their totals from the partials of nblk blocks.  One workgroup; every
// thread folds all components of its blocks in one sweep (block order per
// component, as fold() does), then the block reductions combine the threads.
template <int NM>
__global__ __launch_bounds__(MCM_BLOCK) void k_c_fold(DevState d_, int with_bounds) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    const int nblk = d.nblk;
    Scalars* sc = d.sc;
    constexpr int nm = NM;
    constexpr int nv = 6 + 5 * NM;
    double acc[nv];
#pragma unroll
    for (int k = 0; k < nv; k++) acc[k] = k < 3 ? 1.0e20 : (k < 6 ? -1.0e20 : 0.0);
    for (int b = threadIdx.x; b < nblk; b += MCM_BLOCK) {
#pragma unroll
        for (int k = 0; k < nv; k++) {
            const double x = d.red[(long)k * nblk + b];
            acc[k] = k < 3 ? fmin(acc[k], x) : (k < 6 ? fmax(acc[k], x) : acc[k] + x);
        }
    }
    double tot[nv];
#pragma unroll
    for (int k = 0; k < nv; k++)
        tot[k] = k < 3 ? block_reduce<1>(acc[k]) : (k < 6 ? block_reduce<2>(acc[k]) : block_reduce<0>(acc[k]));
    if (threadIdx.x != 0 || sc->abort) return;
    if (with_bounds)
        for (int a = 0; a < 3; a++) { sc->coord_min[a] = tot[a]; sc->coord_max[a] = tot[3 + a]; }
    double ke = 0.0, ie = 0.0;
    for (int m = 0; m < nm; m++) {
        sc->mat_ke[m] = tot[6 + m];
        sc->mat_ie[m] = tot[6 + nm + m];
        for (int a = 0; a < 3; a++) sc->mat_mom[m][a] = tot[6 + 2 * nm + 3 * m + a];
        ke = ke + tot[6 + m];
        ie = ie + tot[6 + nm + m];
    }
    sc->kinetice = ke;
    sc->internale = ie;
    sc->totale = ke + ie;
    if (with_bounds) sc->timestep += 1;
}

// Fused epilogue of the momentum pass for particle i with acceleration acc:
// mcm_boundcond, the particle's terms of mcm_calctotale (kinetic energy
// from the previous |v|), mcm_update_velocity and (MOVE) mcm_move with the
// symmetry reflection of mcm_check_sym_pen.  Without MOVE (XSPH on) the new
// velocity is stored and k_xsph / k_xsph_move finish the step.
template <bool DIAG, bool MOVE>
__device__ __forceinline__ void integrate_particle(const DevState& d, int i, double acc[3], const double xi[3],
                                                   CPart& cp) {
    Scalars* sc = d.sc;
    if (d.baseaccel)   // mcm_boundcond.f90:23-29
        for (int a = 0; a < 3; a++) acc[a] = acc[a] + d.base_a[a];
    boundcond(d.dispbc[i], acc);
    cp.mi = d.pmat[i];   // (its energy terms: the constitutive pass)
    double v[3] = {d.v[0][i], d.v[1][i], d.v[2][i]};
    // mcm_update_velocity.f90:29-56
    double dtn = 0.5 * (sc->dt + sc->dtold);
    double va = 0.0;
    for (int a = 0; a < 3; a++) {
        v[a] = (d.drelax ? d.drelax_scale * v[a] : v[a]) + acc[a] * dtn;   // :31-41 with dynamic relaxation
        va = va + v[a] * v[a];
    }
    d.vabs[i] = sqrt(va);
    if (!MOVE) {
        for (int a = 0; a < 3; a++) d.v[a][i] = v[a];
        if (DIAG)
            for (int a = 0; a < 3; a++) d.a[a][i] = acc[a];
        return;
    }
    // mcm_move.f90:32-55
    double x[3];
    for (int a = 0; a < 3; a++) {
        x[a] = xi[a] + v[a] * sc->dt;
        cp.x[a] = x[a];
    }
    if (d.boundary) sym_pen(d, x, v);
    for (int a = 0; a < 3; a++) {
        d.v[a][i] = v[a];
        d.x0[a][i] = x[a];
    }
    if (DIAG)
        for (int a = 0; a < 3; a++) d.a[a][i] = acc[a];
}

// Thread-per-particle momentum kernel (CORR & 3 as k_strain_const's CORR;
// CORR & 4: runs with a cap-model material, whose stress is not symmetric).
template <int MODE, int CORR, bool DIAG>
__global__ __launch_bounds__(MCM_BLOCK) __attribute__((amdgpu_waves_per_eu((CORR & 3) == 2 ? MCM_WPE_MOM1 : MCM_WPE_MOM3))) void k_momentum(
    DevState d_, int cur) {
    Blk bk;
    const DevState& d = enter<true>(d_, bk);
    const int lb = bk.b;
    Scalars* sc = d.sc;
    // particles in the search's sorted order (the order of the lists)
    const int t = lb * MCM_BLOCK + threadIdx.x;
    const bool live = t < d.np && !sc->abort;
    const int i = live ? d.sorted[t] : t;
    CPart cp;
    init_c_part(cp);
    constexpr bool ASYM = (CORR & 4) != 0;
    constexpr int DISC = CORR & 3;
    if (live) {
        const MRecT<ASYM> me = load_mrec_t<ASYM>(d, i);
        const double xi[3] = {me.x, me.y, me.z};
        double acc[3];
        // records within the fast-path ranges (flag written by the search,
        // the strain pass and the constitutive pass)
        const bool fast = (d.range[0] & RANGE_RM) == 0;
        if (DISC == 1) {
            if (fast) accel_corr<true, ASYM>(d, i, t, me, acc);
            else accel_corr<false, ASYM>(d, i, t, me, acc);
        } else if (DISC == 2) {
            if (fast) accel_corr_fast<true, ASYM>(d, i, t, me, acc);
            else accel_corr_fast<false, ASYM>(d, i, t, me, acc);
        } else {
            if (fast) accel_plain<true, ASYM>(d, i, t, me, acc);
            else accel_plain<false, ASYM>(d, i, t, me, acc);
        }
        if (MODE == MODE_MOMENTUM) {
            for (int a = 0; a < 3; a++) d.a[a][i] = acc[a];
        } else {
            integrate_particle<DIAG, MODE != MODE_FVEL>(d, i, acc, xi, cp);
        }
    }
    if (MODE == MODE_FUSED) reduce_bound_partials(d, cp, true, lb);   // with XSPH: k_xsph_move
}

// ------------------------------------------------------ optional pair passes
// mcm_rep_contact.f90:20-98 (with mcm_calc_fij, :100-161): Monaghan's
// repulsive acceleration of real particle i from its real contact partners
// (other materials within 2h, listed separately under penalty contact),
// after the constitutive update: x after the strain pass, the new h.
__global__ __launch_bounds__(MCM_BLOCK) void k_rep_cont(DevState d_, int cur) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    const int i = bk.b * MCM_BLOCK + threadIdx.x;
    if (i >= d.np || d.sc->abort) return;
    const int nxt = cur ^ 1;
    double r[3] = {0.0, 0.0, 0.0};
    const int nc = d.ncont[i];
    if (nc > 0) {
        const MatParams& mi = d.mat[d.pmat[i]];
        const double ki = mi.k_cont, en = mi.n_cont;
        const double xi[3] = {d.x1[0][i], d.x1[1][i], d.x1[2][i]};
        const double hi = d.hbuf[nxt][i];
        const int* lst = cont_list(d, i);
        for (int k = 0; k < nc; k++) {
            const int j = lst[(long)k * 64];
            const double hj = d.hbuf[nxt][j], massj = d.mass[j];
            const double kj = d.mat[d.pmat[j]].k_cont;
            const double havg = 0.5 * (hi + hj);
            // inverse of W at the particle spacing (B-spline, :57-66)
            const double ratio = hj / havg;
            double inv_w;
            if (ratio < 1.0) {
                inv_w = 1.0 / (((0.75 * ratio - 1.5) * ratio) * ratio + 1.0);
            } else {
                const double t = 2.0 - ratio;
                inv_w = 1.0 / (t * (t * t));
            }
            double dwdx[3], rad;
            kernel_grad(xi[0] - d.x1[0][j], xi[1] - d.x1[1][j], xi[2] - d.x1[2][j], havg, d.pi, dwdx, &rad);
            const double z = rad / havg;
            double fij;
            if (z < 1.0) {
                fij = inv_w * (((0.75 * z - 1.5) * z) * z + 1.0);
            } else if (z < 2.0) {
                const double t = 2.0 - z;
                fij = 0.25 * inv_w * (t * (t * t));
            } else {
                fij = 0.0;
            }
            const double kij = ki + kj;
            for (int l = 0; l < 3; l++) r[l] = r[l] - massj * (kij * pow(fij, en)) * dwdx[l];
        }
    }
    for (int l = 0; l < 3; l++) d.rep[l][i] = r[l];
}

// XSPH smoothing of mcm_update_velocity.f90:61-81 for real particle i, from
// the new velocities of its list (reals, self, then ghosts with their
// start-of-step copies): sum of m_j / (rho_j + rho_i)/2 (v_j - v_i) W(x_i - x_j,
// h_j) into sv.
__global__ __launch_bounds__(MCM_BLOCK) void k_xsph(DevState d_, int cur) {
    Blk bk;
    const DevState& d = enter<true>(d_, bk);
    const int i = bk.b * MCM_BLOCK + threadIdx.x;
    if (i >= d.np || d.sc->abort) return;
    const int nxt = cur ^ 1;
    const double xi[3] = {d.x1[0][i], d.x1[1][i], d.x1[2][i]};
    const double vi[3] = {d.v[0][i], d.v[1][i], d.v[2][i]};
    const double rhoi = d.rbuf[nxt][i];
    double s[3] = {0.0, 0.0, 0.0};
    const int n = d.ntot[i];
    const int* lst = nbr_row(d, d.spos[i]);
    for (int k = 0; k < n; k++) {
        const int u = lst[(long)k * 64];
        double rad = 0.0;
        for (int a = 0; a < 3; a++) {
            const double xij = xi[a] - d.x1[a][u];
            rad = rad + xij * xij;
        }
        rad = sqrt(rad);
        const double w = kernel_w(rad, d.hbuf[nxt][u], d.pi);
        const double volj = d.mass[u] / (0.5 * (d.rbuf[nxt][u] + rhoi));
        for (int a = 0; a < 3; a++) s[a] = s[a] + volj * (d.v[a][u] - vi[a]) * w;
    }
    for (int a = 0; a < 3; a++) d.sv[a][i] = s[a];
}

// v += 0.1 sv once every particle's smoothing is known (:83-87)
__global__ __launch_bounds__(MCM_BLOCK) void k_xsph_apply(DevState d) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    if (i >= d.np || d.sc->abort) return;
    for (int a = 0; a < 3; a++) d.v[a][i] = d.v[a][i] + 0.100 * d.sv[a][i];
}

// Fused step with XSPH: the smoothed velocity, then mcm_move and the
// symmetry reflection; coordinate-bound partials of the block (its energy
// partials were written by the momentum kernel).
__global__ __launch_bounds__(MCM_BLOCK) void k_xsph_move(DevState d_, int cur) {
    Blk bk;
    const DevState& d = enter<true>(d_, bk);
    const int lb = bk.b;
    const int i = lb * MCM_BLOCK + threadIdx.x;
    Scalars* sc = d.sc;
    const bool live = i < d.np && !sc->abort;
    double xb[3] = {0.0, 0.0, 0.0};   // bounds before the reflection, as mcm_move.f90:40-55
    if (live) {
        double v[3], x[3];
        for (int a = 0; a < 3; a++) v[a] = d.v[a][i] + 0.100 * d.sv[a][i];
        for (int a = 0; a < 3; a++) {
            x[a] = d.x1[a][i] + v[a] * sc->dt;
            xb[a] = x[a];
        }
        if (d.boundary) sym_pen(d, x, v);
        for (int a = 0; a < 3; a++) {
            d.v[a][i] = v[a];
            d.x0[a][i] = x[a];
        }
    }
    for (int a = 0; a < 3; a++) {
        const double lo = block_reduce<1>(live ? xb[a] : 1.0e20);
        const double hi = block_reduce<2>(live ? xb[a] : -1.0e20);
        if (threadIdx.x == 0) { *red_at(d, lb, a) = lo; *red_at(d, lb, 3 + a) = hi; }
    }
}

// ---------------------------------------------- granular integration stages
__global__ __launch_bounds__(MCM_BLOCK) void k_boundcond(DevState d) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    if (i >= d.np || d.sc->abort) return;
    double acc[3] = {d.a[0][i], d.a[1][i], d.a[2][i]};
    if (d.baseaccel)   // mcm_boundcond.f90:23-29
        for (int a = 0; a < 3; a++) acc[a] = acc[a] + d.base_a[a];
    boundcond(d.dispbc[i], acc);
    for (int a = 0; a < 3; a++) d.a[a][i] = acc[a];
}

__global__ __launch_bounds__(MCM_BLOCK) void k_calctotale(DevState d) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    CPart cp;
    init_c_part(cp);
    if (i < d.np && !d.sc->abort) {
        const int mi = d.pmat[i];
        const double mass = d.mass[i], vabs = d.vabs[i];
        double v[3] = {d.v[0][i], d.v[1][i], d.v[2][i]};
        energy_terms(cp, mi, mass, vabs, d.e[i], v);
    }
    reduce_c_partials(d, cp, false, blockIdx.x);
}

__global__ __launch_bounds__(MCM_BLOCK) void k_update_velocity(DevState d) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    Scalars* sc = d.sc;
    if (i >= d.np || sc->abort) return;
    double dtn = 0.5 * (sc->dt + sc->dtold);
    double va = 0.0;
    for (int a = 0; a < 3; a++) {
        double v = (d.drelax ? d.drelax_scale * d.v[a][i] : d.v[a][i]) + d.a[a][i] * dtn;
        d.v[a][i] = v;
        va = va + v * v;
    }
    d.vabs[i] = sqrt(va);
}

__global__ __launch_bounds__(MCM_BLOCK) void k_move(DevState d) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    Scalars* sc = d.sc;
    double vals[6] = {1.0e20, 1.0e20, 1.0e20, -1.0e20, -1.0e20, -1.0e20};
    if (i < d.np && !sc->abort) {
        double x[3], v[3];
        for (int a = 0; a < 3; a++) {
            v[a] = d.v[a][i];
            x[a] = d.x1[a][i] + v[a] * sc->dt;
            vals[a] = x[a];
            vals[3 + a] = x[a];
        }
        if (d.boundary) sym_pen(d, x, v);
        for (int a = 0; a < 3; a++) { d.v[a][i] = v[a]; d.x0[a][i] = x[a]; }
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double r = k < 3 ? block_reduce<1>(vals[k]) : block_reduce<2>(vals[k]);
        if (threadIdx.x == 0) d.red[(long)blockIdx.x * 6 + k] = r;
    }
}

__global__ __launch_bounds__(MCM_BLOCK) void k_move_fold(DevState d, int nblk) {
    Scalars* sc = d.sc;
    double r[6];
#pragma unroll
    for (int k = 0; k < 6; k++) r[k] = k < 3 ? fold<1>(d.red, 6, k, nblk, 1.0e20) : fold<2>(d.red, 6, k, nblk, -1.0e20);
    if (threadIdx.x == 0 && !sc->abort) {
        for (int k = 0; k < 3; k++) { sc->coord_min[k] = r[k]; sc->coord_max[k] = r[3 + k]; }
        sc->timestep += 1;
    }
}

// mcm_calccritts.f90:32-87 over the current state (initialisation path)
__global__ __launch_bounds__(MCM_BLOCK) void k_calccritts(DevState d, int cur) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    Scalars* sc = d.sc;
    double crit = 1.0e20;
    if (i < d.np && !sc->abort) {
        PState s;
        load_state(d, i, cur, s);
        s.tracerod = d.tracerod[i];
        const MatParams& m = d.mat[d.pmat[i]];
        s.critts = critts(s, m, d.tcrit_opt, d.mindist[i]);
        d.critts_p[i] = s.critts;
        crit = s.critts;
    }
    double cb = block_reduce<1>(crit);
    if (threadIdx.x == 0) d.red[blockIdx.x] = cb;
}

__global__ __launch_bounds__(MCM_BLOCK) void k_calccritts_fold(DevState d, int nblk) {
    double c1 = fold<1>(d.red, 1, 0, nblk, 1.0e20);
    if (threadIdx.x == 0) d.sc->critts = c1;
}

// ------------------------------------------------------------------ output
// The fields the reference's writers read, packed per particle for one
// device-to-host copy (mcm_gpu_output_fields): x(3), v(3), p, sigma 11 22 33
// 12 13 23, e, efps, rho, h, converted to T.  x is the position after the
// strain pass between the strain and move stages (xsel), rho and h those of
// buffer hb.
#define OUT_WORDS 17
template <typename T>
__global__ __launch_bounds__(MCM_BLOCK) void k_pack_output(DevState d, int hb, int xsel, T* out) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    if (i >= d.np) return;
    double* const* x = xsel ? d.x1 : d.x0;
    T* r = out + (long)i * OUT_WORDS;
    for (int a = 0; a < 3; a++) {
        r[a] = (T)x[a][i];
        r[3 + a] = (T)d.v[a][i];
    }
    r[6] = (T)d.p[i];
    r[7] = (T)d.sig[S11][i];
    r[8] = (T)d.sig[S22][i];
    r[9] = (T)d.sig[S33][i];
    r[10] = (T)d.sig[S12][i];
    r[11] = (T)d.sig[S13][i];
    r[12] = (T)d.sig[S23][i];
    r[13] = (T)d.e[i];
    r[14] = (T)d.efps[i];
    r[15] = (T)d.rbuf[hb][i];
    r[16] = (T)d.hbuf[hb][i];
}
