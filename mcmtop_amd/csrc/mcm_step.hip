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

enum { MODE_FUSED = 0, MODE_STRAIN = 1, MODE_CONST = 2, MODE_MOMENTUM = 1 };

// ------------------------------------------------------------- reductions
__device__ __forceinline__ double vload(const double* p) { return *(volatile const double*)p; }
__device__ __forceinline__ double wave_min(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}

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

// The last block of a grid folds the per-block partials red[b*stride + k]
// of all blocks: thread t walks blocks t, t+BLOCK, ... in order, then the
// block reduction combines the threads in a fixed order.
template <int OP>
__device__ __forceinline__ double fold_partials(const double* red, int stride, int k, double init) {
    double v = init;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += MCM_BLOCK) {
        double x = vload(&red[(long)b * stride + k]);
        v = OP == 0 ? v + x : (OP == 1 ? fmin(v, x) : fmax(v, x));
    }
    return block_reduce<OP>(v);
}

__device__ __forceinline__ bool last_block(unsigned int* counter) {
    __shared__ int s_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int t = atomicAdd(counter, 1u);
        s_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    return s_last != 0;
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
// Neighbour records: the strain passes read rs[u] = {x, y, z, vx, vy, vz,
// m/rho, h} (8 doubles, 64 B) and the momentum passes rm[u] = {x, y, z, h,
// m, m/rho, A11, A22, A33, A12, A13, A23} (12 doubles, 96 B), one aligned
// record per neighbour instead of one load per state array.
struct SRec { double x, y, z, vx, vy, vz, vol, h; };
struct MRec { double x, y, z, h, m, vol, A[6]; };

__device__ __forceinline__ SRec load_srec(const double* rs, int j) {
    const double2* r = reinterpret_cast<const double2*>(rs) + 4 * (long)j;
    double2 a = r[0], b = r[1], c = r[2], e = r[3];
    return SRec{a.x, a.y, b.x, b.y, c.x, c.y, e.x, e.y};
}
__device__ __forceinline__ MRec load_mrec(const double* rm, int j) {
    const double2* r = reinterpret_cast<const double2*>(rm) + 6 * (long)j;
    double2 a = r[0], b = r[1], c = r[2], e = r[3], f = r[4], g = r[5];
    return MRec{a.x, a.y, b.x, b.y, c.x, c.y, {e.x, e.y, f.x, f.y, g.x, g.y}};
}

// plain strain rate, mcm_oldstrain.f90:34-64, positions half a step back
// for real particles only (mcm_strain.f90:27-31)
__device__ __forceinline__ void strain_plain(const DevState& d, int i, double dt, const double xi[3],
                                             const double vi[3], double hi, double g[3][3]) {
    const int n = d.ntot[i];
    const int* lst = d.nbr + i;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) g[a][b] = 0.0;
    for (int k = 0; k < n; k++) {
        int j = lst[(long)k * d.npad];
        SRec r = load_srec(d.rs, j);
        double sh = j < d.np ? dt : 0.0;
        double xj0 = r.x - 0.5 * r.vx * sh;
        double xj1 = r.y - 0.5 * r.vy * sh;
        double xj2 = r.z - 0.5 * r.vz * sh;
        double havg = 0.5 * (hi + r.h);
        double dwdx[3], rad;
        kernel_grad(xi[0] - xj0, xi[1] - xj1, xi[2] - xj2, havg, d.pi, dwdx, &rad);
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
__device__ __forceinline__ void corr_pair(const SRec& r, double sh, const double xi[3], double hi, double pi,
                                          double xj[3], double& w, double dwdx[3]) {
    xj[0] = r.x - 0.5 * r.vx * sh;
    xj[1] = r.y - 0.5 * r.vy * sh;
    xj[2] = r.z - 0.5 * r.vz * sh;
    double havg = 0.5 * (hi + r.h);
    double rad;
    kernel_grad(xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2], havg, pi, dwdx, &rad);
    w = kernel_w(rad, havg, pi);
}

// corrected strain rate, mcm_corr_strain.f90:37-146, in the reference's three
// passes and evaluation order
__device__ __forceinline__ void strain_corr(const DevState& d, int i, double dt, const double xi[3], double hi,
                                            double f[3][3]) {
    const int n = d.ntot[i];
    const int* lst = d.nbr + i;
    double sw = 0.0, sg[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < n; k++) {
        int j = lst[(long)k * d.npad];
        SRec r = load_srec(d.rs, j);
        double xj[3], w, dwdx[3];
        corr_pair(r, j < d.np ? dt : 0.0, xi, hi, d.pi, xj, w, dwdx);
        sw = sw + r.vol * w;
        for (int l = 0; l < 3; l++) sg[l] = sg[l] + r.vol * dwdx[l];
    }
    const double sw2 = sw * sw;
    double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, invB[9];
    for (int k = 0; k < n; k++) {
        int j = lst[(long)k * d.npad];
        SRec r = load_srec(d.rs, j);
        double xj[3], w, dwdx[3], cg[3];
        corr_pair(r, j < d.np ? dt : 0.0, xi, hi, d.pi, xj, w, dwdx);
        for (int m = 0; m < 3; m++) cg[m] = (dwdx[m] * sw - w * sg[m]) / sw2;
        for (int m = 0; m < 3; m++)
            for (int nn = 0; nn < 3; nn++) B[nn + 3 * m] = B[nn + 3 * m] + r.vol * (xj[nn] - xi[nn]) * cg[m];
    }
    inverse3(B, invB);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) f[a][b] = 0.0;
    for (int k = 0; k < n; k++) {
        int j = lst[(long)k * d.npad];
        SRec r = load_srec(d.rs, j);
        double xj[3], w, dwdx[3], cg[3], g[3];
        corr_pair(r, j < d.np ? dt : 0.0, xi, hi, d.pi, xj, w, dwdx);
        for (int m = 0; m < 3; m++) cg[m] = (dwdx[m] * sw - w * sg[m]) / sw2;
        for (int q = 0; q < 3; q++) g[q] = invB[q] * cg[0] + invB[q + 3] * cg[1] + invB[q + 6] * cg[2];
        double vj[3] = {r.vx, r.vy, r.vz};
        for (int nn = 0; nn < 3; nn++)
            for (int m = 0; m < 3; m++) f[nn][m] = f[nn][m] + r.vol * vj[nn] * g[m];
    }
}

// corrected strain rate in one pass: the correction is linear in the pass-1
// sums, so B and the contracted gradient follow from single-pass moments
// (same result up to rounding; the fast variant)
__device__ __forceinline__ void strain_corr_fast(const DevState& d, int i, double dt, const double xi[3],
                                                 double hi, double f[3][3]) {
    const int n = d.ntot[i];
    const int* lst = d.nbr + i;
    double sw = 0.0, sg[3] = {0, 0, 0}, P[3][3] = {{0}}, Pw[3] = {0, 0, 0}, Q[3][3] = {{0}}, Qw[3] = {0, 0, 0};
    for (int k = 0; k < n; k++) {
        int j = lst[(long)k * d.npad];
        SRec r = load_srec(d.rs, j);
        double xj[3], w, dwdx[3];
        corr_pair(r, j < d.np ? dt : 0.0, xi, hi, d.pi, xj, w, dwdx);
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
// CORR: 0 plain SPH, 1 corrected (reference order), 2 corrected (one pass)
template <int MODE, int CORR, bool DIAG>
__global__ __launch_bounds__(MCM_BLOCK) void k_strain_const(DevState d, int cur) {
    const int lb = xcd_block();
    const int i = lb * MCM_BLOCK + threadIdx.x;
    const int nxt = cur ^ 1;
    Scalars* sc = d.sc;
    const bool live = i < d.np && !sc->abort;
    const double dt = sc->dt;
    double crit = 1.0e20, hm = 0.0;
    int eos_err = 0;
    if (live) {
        PState s;
        double x1z_keep = 0.0;
        if (MODE != MODE_CONST) {
            double xr[3], vi[3], xi[3];
            for (int a = 0; a < 3; a++) {
                xr[a] = d.x0[a][i];
                vi[a] = d.v[a][i];
                xi[a] = xr[a] - 0.5 * vi[a] * dt;
            }
            double hi = d.hbuf[cur][i];
            double g[3][3];
            if (CORR == 1) strain_corr(d, i, dt, xi, hi, g);
            else if (CORR == 2) strain_corr_fast(d, i, dt, xi, hi, g);
            else strain_plain(d, i, dt, xi, vi, hi, g);
            rod_spin(g, s);
            // the reference moves x back by the same half step: (x - t) + t
            double x1[3];
            for (int a = 0; a < 3; a++) {
                x1[a] = xi[a] + 0.5 * vi[a] * dt;
                d.x1[a][i] = x1[a];
            }
            x1z_keep = x1[2];
            double2* rm = reinterpret_cast<double2*>(d.rm) + 6 * (long)i;
            rm[0] = make_double2(x1[0], x1[1]);
            if (MODE == MODE_STRAIN) rm[1] = make_double2(x1[2], hi);
            if (MODE == MODE_STRAIN) {
                for (int k = 0; k < 6; k++) d.rod[k][i] = s.rod[k];
                for (int k = 0; k < 3; k++) d.spin[k][i] = s.spin[k];
                d.tracerod[i] = s.tracerod;
                d.hbuf[nxt][i] = hi;
                d.rbuf[nxt][i] = d.rbuf[cur][i];
            }
        } else {
            for (int k = 0; k < 6; k++) s.rod[k] = d.rod[k][i];
            for (int k = 0; k < 3; k++) s.spin[k] = d.spin[k][i];
            s.tracerod = d.tracerod[i];
        }
        if (MODE != MODE_STRAIN) {
            // granular mode: h and rho were staged in the next buffers by the
            // strain and update_h stages
            load_state(d, i, MODE == MODE_CONST ? nxt : cur, s);
            const int mi = d.pmat[i];
            const MatParams& m = d.mat[mi];
            if (MODE == MODE_FUSED && d.h_opt > 0) {
                if (update_h(s, d.h_opt, dt)) hm = s.h;
            }
            rhoupd(s, m, dt);
            if (m.model == 3) for (int k = 0; k < 6; k++) s.alfa[k] = d.alfa[k][i];
            if (m.model == 9) for (int k = 0; k < 6; k++) s.s[k] = d.sd[k][i];
            double defps_prev = d.defps_prev[i];
            int active = constitutive_one(s, m, mi + 1, dt, defps_prev, &eos_err);
            s.critts = critts(s, m, d.tcrit_opt, d.mindist[i]);
            crit = s.critts;
            // write back
            d.rbuf[nxt][i] = s.rho;
            d.hbuf[nxt][i] = s.h;
            if (d.h_opt > 0) d.hold[i] = s.hold;
            for (int k = 0; k < 6; k++) d.sig[k][i] = s.sig[k];
            if (m.model == 3) for (int k = 0; k < 6; k++) d.alfa[k][i] = s.alfa[k];
            if (m.model == 9 || (DIAG && m.model == 10)) for (int k = 0; k < 6; k++) d.sd[k][i] = s.s[k];
            if (DIAG && m.model == 3) for (int k = 0; k < 6; k++) d.sd[k][i] = s.s[k];
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
            double2* rm = reinterpret_cast<double2*>(d.rm) + 6 * (long)i;
            double x1z = MODE == MODE_CONST ? d.x1[2][i] : x1z_keep;
            rm[1] = make_double2(x1z, s.h);
            rm[2] = make_double2(s.mass, s.mass / s.rho);
            rm[3] = make_double2((s.sig[S11] - s.q) / r2, (s.sig[S22] - s.q) / r2);
            rm[4] = make_double2((s.sig[S33] - s.q) / r2, (s.sig[S12] - 0.0) / r2);
            rm[5] = make_double2((s.sig[S13] - 0.0) / r2, (s.sig[S23] - 0.0) / r2);
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
    }
    if (MODE == MODE_STRAIN) return;
    if (eos_err) atomicOr(&sc->error_flags, 16);
    // block min of critts, max of updated h -> partials -> last block
    double cb = block_reduce<1>(crit);
    double hb = block_reduce<2>(hm);
    if (threadIdx.x == 0) {
        d.red[2 * lb] = cb;
        d.red[2 * lb + 1] = hb;
    }
    if (!last_block(&sc->done_b)) return;
    double c1 = fold_partials<1>(d.red, 2, 0, 1.0e20);
    double h1 = fold_partials<2>(d.red, 2, 1, 0.0);
    if (threadIdx.x == 0) {
        if (!sc->abort) {
            sc->critts = c1;
            if (MODE == MODE_FUSED && d.h_opt > 0) sc->hmax = h1;
            if (MODE == MODE_FUSED) {   // mcm_solution.f90:58-62
                sc->dtold = sc->dt;
                sc->dt = d.tssf * sc->critts;
                sc->dt = fmin(sc->dt, 1.1 * sc->dtold);
                sc->ptime = sc->ptime + sc->dt;
            }
        }
        sc->done_b = 0;
    }
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
    if (!last_block(&sc->done_r)) return;
    double h1 = fold_partials<2>(d.red, 1, 0, 0.0);
    if (threadIdx.x == 0) {
        sc->hmax = h1;
        sc->done_r = 0;
    }
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
__global__ __launch_bounds__(MCM_BLOCK) void k_ghost_stress(DevState d, int cur) {
    const int nxt = cur ^ 1;
    if (d.sc->abort) return;
    const int ngp = d.sc->ngp;
    for (int g = blockIdx.x * MCM_BLOCK + threadIdx.x; g < ngp; g += gridDim.x * MCM_BLOCK) {
        const int u = d.np + g, p = d.gpar[g];
        double rg = d.rbuf[nxt][u];
        double r2 = rg * rg;
        double q = d.q[p];
        double2* rm = reinterpret_cast<double2*>(d.rm) + 6 * (long)u;
        rm[3] = make_double2((d.sig[S11][p] - q) / r2, (d.sig[S22][p] - q) / r2);
        rm[4] = make_double2((d.sig[S33][p] - q) / r2, -d.sig[S12][p] / r2);
        rm[5] = make_double2(-d.sig[S13][p] / r2, -d.sig[S23][p] / r2);
    }
}

// ------------------------------------------------------------- momentum
__device__ __forceinline__ int sym_index(int r, int c) {
    return r == c ? r : ((r + c == 1) ? S12 : (r + c == 2 ? S13 : S23));
}

// mcm_oldacceleration.f90:60-89
__device__ __forceinline__ void accel_plain(const DevState& d, int i, const MRec& me, double acc[3]) {
    const int n = d.ntot[i];
    const int* lst = d.nbr + i;
    acc[0] = 0.0; acc[1] = 0.0; acc[2] = 0.0;
    for (int k = 0; k < n; k++) {
        int j = lst[(long)k * d.npad];
        MRec r = load_mrec(d.rm, j);
        double havg = 0.5 * (me.h + r.h);
        double dwdx[3], rad;
        kernel_grad(me.x - r.x, me.y - r.y, me.z - r.z, havg, d.pi, dwdx, &rad);
        for (int nn = 0; nn < 3; nn++) {
            double ds = 0.0;
            for (int mm = 0; mm < 3; mm++) {
                int c = sym_index(mm, nn);
                ds = ds + (r.A[c] + me.A[c]) * dwdx[mm];
            }
            acc[nn] = acc[nn] + r.m * ds;
        }
    }
}

// kernel value and gradient centred at the neighbour: W(x_j, x_i),
// grad_j W (mcm_corr_acceleration.f90:73-74)
__device__ __forceinline__ void corr_pair_j(const MRec& r, const MRec& me, double pi, double& w, double dwdx[3]) {
    double havg = 0.5 * (me.h + r.h);
    double rad;
    kernel_grad(r.x - me.x, r.y - me.y, r.z - me.z, havg, pi, dwdx, &rad);
    w = kernel_w(rad, havg, pi);
}

// mcm_corr_acceleration.f90:37-163 in the reference's three passes: the self
// slot (k == nnbr) contributes nothing to deltasig, the mass of the last
// neighbour multiplies the sum and B^-1 enters transposed
__device__ __forceinline__ void accel_corr(const DevState& d, int i, const MRec& me, double acc[3]) {
    const int n = d.ntot[i];
    const int self_k = d.nreal[i] - 1;
    const int* lst = d.nbr + i;
    double sw = 0.0, sg[3] = {0.0, 0.0, 0.0}, massj = 0.0;
    for (int k = 0; k < n; k++) {
        MRec r = load_mrec(d.rm, lst[(long)k * d.npad]);
        double w, dwdx[3];
        corr_pair_j(r, me, d.pi, w, dwdx);
        sw = sw + r.vol * w;
        for (int l = 0; l < 3; l++) sg[l] = sg[l] + r.vol * dwdx[l];
        massj = r.m;
    }
    const double sw2 = sw * sw;
    double B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, invB[9];
    for (int k = 0; k < n; k++) {
        MRec r = load_mrec(d.rm, lst[(long)k * d.npad]);
        double w, dwdx[3], cg[3];
        corr_pair_j(r, me, d.pi, w, dwdx);
        for (int m = 0; m < 3; m++) cg[m] = (dwdx[m] * sw - w * sg[m]) / sw2;
        double dx[3] = {r.x - me.x, r.y - me.y, r.z - me.z};
        for (int m = 0; m < 3; m++)
            for (int nn = 0; nn < 3; nn++) B[nn + 3 * m] = B[nn + 3 * m] + r.vol * dx[nn] * cg[m];
    }
    inverse3(B, invB);
    // deltasig(m,n,l) is symmetric in (m,n): 6 x 3 accumulators
    double D[6][3] = {{0.0}};
    for (int k = 0; k < n; k++) {
        if (k == self_k) continue;   // dsig = 0 there
        MRec r = load_mrec(d.rm, lst[(long)k * d.npad]);
        double w, dwdx[3], cg[3];
        corr_pair_j(r, me, d.pi, w, dwdx);
        for (int m = 0; m < 3; m++) cg[m] = (dwdx[m] * sw - w * sg[m]) / sw2;
        for (int c = 0; c < 6; c++) {
            double dsig = r.A[c] + me.A[c];
            D[c][0] = D[c][0] + dsig * cg[0];
            D[c][1] = D[c][1] + dsig * cg[1];
            D[c][2] = D[c][2] + dsig * cg[2];
        }
    }
    for (int nn = 0; nn < 3; nn++) {
        double a = 0.0;
        for (int mm = 0; mm < 3; mm++) {
            int c = sym_index(nn, mm);
            for (int l = 0; l < 3; l++) a = a + massj * D[c][l] * invB[l + 3 * mm];
        }
        acc[nn] = a;
    }
}

// corrected acceleration in one pass from single-pass moments (fast variant)
__device__ __forceinline__ void accel_corr_fast(const DevState& d, int i, const MRec& me, double acc[3]) {
    const int n = d.ntot[i];
    const int self_k = d.nreal[i] - 1;
    const int* lst = d.nbr + i;
    double sw = 0.0, sg[3] = {0, 0, 0}, P[3][3] = {{0}}, Pw[3] = {0, 0, 0};
    double E1[6][3] = {{0}}, E2[6] = {0, 0, 0, 0, 0, 0};
    double mlast = 0.0;
    for (int k = 0; k < n; k++) {
        MRec r = load_mrec(d.rm, lst[(long)k * d.npad]);
        double w, dwdx[3];
        corr_pair_j(r, me, d.pi, w, dwdx);
        double dx[3] = {r.x - me.x, r.y - me.y, r.z - me.z};
        double vw = r.vol * w;
        sw += vw;
        for (int a = 0; a < 3; a++) {
            sg[a] += r.vol * dwdx[a];
            Pw[a] += dx[a] * vw;
            for (int b = 0; b < 3; b++) P[a][b] += dx[a] * (r.vol * dwdx[b]);
        }
        if (k != self_k) {
            for (int c = 0; c < 6; c++) {
                double ds = r.A[c] + me.A[c];
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
            int c = sym_index(nn, mm);
            for (int l = 0; l < 3; l++) sacc += (E1[c][l] * sw - E2[c] * sg[l]) / sw2 * invB[l + 3 * mm];
        }
        acc[nn] = mlast * sacc;
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

// energy and coordinate-bound partials: [cmin 3][cmax 3][ke nm][ie nm][mom 3nm]
#define RED_C (6 + 5 * MCM_MAX_MAT)

__device__ void reduce_c_partials(const DevState& d, const double (&vals)[RED_C], bool finalize_bounds, int lb) {
    Scalars* sc = d.sc;
    const int nm = d.nummat;
    const int nv = 6 + 5 * nm;
    double* out = d.red + (long)lb * RED_C;
#pragma unroll
    for (int k = 0; k < RED_C; k++) {
        if (k < nv) {
            double r = k < 3 ? block_reduce<1>(vals[k]) : (k < 6 ? block_reduce<2>(vals[k]) : block_reduce<0>(vals[k]));
            if (threadIdx.x == 0) out[k] = r;
        }
    }
    if (!last_block(&sc->done_c)) return;
    double tot[RED_C];
#pragma unroll
    for (int k = 0; k < RED_C; k++) {
        tot[k] = 0.0;
        if (k < nv)
            tot[k] = k < 3 ? fold_partials<1>(d.red, RED_C, k, 1.0e20)
                           : (k < 6 ? fold_partials<2>(d.red, RED_C, k, -1.0e20) : fold_partials<0>(d.red, RED_C, k, 0.0));
    }
    if (threadIdx.x == 0) {
        if (finalize_bounds && !sc->abort) {
            for (int k = 0; k < 3; k++) { sc->coord_min[k] = tot[k]; sc->coord_max[k] = tot[3 + k]; }
        }
        double ke = 0.0, ie = 0.0;
        for (int m = 0; m < MCM_MAX_MAT; m++) {
            if (m >= nm || sc->abort) break;
            sc->mat_ke[m] = tot[6 + m];
            sc->mat_ie[m] = tot[6 + nm + m];
            for (int a = 0; a < 3; a++) sc->mat_mom[m][a] = tot[6 + 2 * nm + 3 * m + a];
            ke = ke + sc->mat_ke[m];
            ie = ie + sc->mat_ie[m];
        }
        if (!sc->abort) {
            sc->kinetice = ke;
            sc->internale = ie;
            sc->totale = ke + ie;
            if (finalize_bounds) sc->timestep += 1;
        }
        sc->done_c = 0;
    }
}

// per-particle energy contributions into the partial vector (mcm_calctotale.f90:33-51)
__device__ __forceinline__ void energy_terms(double (&vals)[RED_C], int nm, int mi, double mass, double vabs,
                                             double e, const double v[3]) {
#pragma unroll
    for (int m = 0; m < MCM_MAX_MAT; m++) {
        if (m == mi) {
            vals[6 + m] = 0.5 * mass * (vabs * vabs);
        }
    }
    // ie and mom slots depend on nm; nm <= MCM_MAX_MAT
#pragma unroll
    for (int k = 6; k < RED_C; k++) {
        int r = k - 6;
        if (r >= nm && r < 2 * nm && r - nm == mi) vals[k] = e;
        if (r >= 2 * nm && r < 5 * nm) {
            int q = r - 2 * nm;
            if (q / 3 == mi) vals[k] = mass * fabs(v[q % 3]);
        }
    }
}

template <int MODE, int CORR, bool DIAG>
__global__ __launch_bounds__(MCM_BLOCK) void k_momentum(DevState d, int cur) {
    const int lb = xcd_block();
    const int i = lb * MCM_BLOCK + threadIdx.x;
    Scalars* sc = d.sc;
    const bool live = i < d.np && !sc->abort;
    double vals[RED_C];
    for (int k = 0; k < RED_C; k++) vals[k] = k < 3 ? 1.0e20 : (k < 6 ? -1.0e20 : 0.0);
    if (live) {
        const MRec me = load_mrec(d.rm, i);
        const double xi[3] = {me.x, me.y, me.z};
        double acc[3];
        if (CORR == 1) accel_corr(d, i, me, acc);
        else if (CORR == 2) accel_corr_fast(d, i, me, acc);
        else accel_plain(d, i, me, acc);
        if (MODE == MODE_MOMENTUM) {
            for (int a = 0; a < 3; a++) d.a[a][i] = acc[a];
        } else {
            boundcond(d.dispbc[i], acc);
            const int mi = d.pmat[i];
            const double mass = d.mass[i];
            double v[3] = {d.v[0][i], d.v[1][i], d.v[2][i]};
            double vabs = d.vabs[i];
            // mcm_calctotale.f90:33-51 (kinetic energy from the previous |v|)
            energy_terms(vals, d.nummat, mi, mass, vabs, d.e[i], v);
            // mcm_update_velocity.f90:29-56
            double dtn = 0.5 * (sc->dt + sc->dtold);
            double va = 0.0;
            for (int a = 0; a < 3; a++) {
                v[a] = v[a] + acc[a] * dtn;
                va = va + v[a] * v[a];
            }
            d.vabs[i] = sqrt(va);
            // mcm_move.f90:32-55
            double x[3];
            for (int a = 0; a < 3; a++) {
                x[a] = xi[a] + v[a] * sc->dt;
                vals[a] = x[a];
                vals[3 + a] = x[a];
            }
            if (d.boundary) sym_pen(d, x, v);
            for (int a = 0; a < 3; a++) {
                d.v[a][i] = v[a];
                d.x0[a][i] = x[a];
            }
            if (DIAG)
                for (int a = 0; a < 3; a++) d.a[a][i] = acc[a];
        }
    }
    if (MODE == MODE_FUSED) reduce_c_partials(d, vals, true, lb);
}

// ---------------------------------------------- granular integration stages
__global__ __launch_bounds__(MCM_BLOCK) void k_boundcond(DevState d) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    if (i >= d.np || d.sc->abort) return;
    double acc[3] = {d.a[0][i], d.a[1][i], d.a[2][i]};
    boundcond(d.dispbc[i], acc);
    for (int a = 0; a < 3; a++) d.a[a][i] = acc[a];
}

__global__ __launch_bounds__(MCM_BLOCK) void k_calctotale(DevState d) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    double vals[RED_C];
    for (int k = 0; k < RED_C; k++) vals[k] = k < 3 ? 1.0e20 : (k < 6 ? -1.0e20 : 0.0);
    if (i < d.np && !d.sc->abort) {
        const int mi = d.pmat[i];
        const double mass = d.mass[i], vabs = d.vabs[i];
        double v[3] = {d.v[0][i], d.v[1][i], d.v[2][i]};
        energy_terms(vals, d.nummat, mi, mass, vabs, d.e[i], v);
    }
    reduce_c_partials(d, vals, false, blockIdx.x);
}

__global__ __launch_bounds__(MCM_BLOCK) void k_update_velocity(DevState d) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    Scalars* sc = d.sc;
    if (i >= d.np || sc->abort) return;
    double dtn = 0.5 * (sc->dt + sc->dtold);
    double va = 0.0;
    for (int a = 0; a < 3; a++) {
        double v = d.v[a][i] + d.a[a][i] * dtn;
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
    if (!last_block(&sc->done_r)) return;
    double r[6];
#pragma unroll
    for (int k = 0; k < 6; k++) r[k] = k < 3 ? fold_partials<1>(d.red, 6, k, 1.0e20) : fold_partials<2>(d.red, 6, k, -1.0e20);
    if (threadIdx.x == 0) {
        if (!sc->abort) {
            for (int k = 0; k < 3; k++) { sc->coord_min[k] = r[k]; sc->coord_max[k] = r[3 + k]; }
            sc->timestep += 1;
        }
        sc->done_r = 0;
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
    if (!last_block(&sc->done_r)) return;
    double c1 = fold_partials<1>(d.red, 1, 0, 1.0e20);
    if (threadIdx.x == 0) {
        sc->critts = c1;
        sc->done_r = 0;
    }
}
