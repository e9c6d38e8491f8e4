// mcm_adjoint.hip -- reverse-mode derivative of the plain-SPH MCM step on gfx950.
//
// A forward run records a state snapshot every k steps (k = 1: the
// recording runs the granular stages and keeps every step's start state and
// intermediates, and the reverse sweep recomputes nothing).  The reverse sweep
// walks the segments backwards.  Each segment is recomputed from its
// checkpoint through the granular stage kernels, which leave every
// intermediate the reverse needs on the device (lists, ghosts, gather
// records, rod/spin, per-particle critts, acceleration); the start-of-step
// state and those intermediates of every step of the segment are kept in
// HBM.  Then, for each step from last to first, its intermediates are put
// back and the adjoint kernels map the adjoint of the end-of-step state to
// the adjoint of the start-of-step state, in the reverse order of the
// stages: move and velocity update, dt update, momentum pair gather,
// constitutive update, strain pair gather.
//
// Pair stages: the real-particle neighbour relation is symmetric (the list
// criterion |x_i - x_j|^2 < (h_i + h_j)^2 at x^n is), so the contribution of
// pair (i, j) of i's sum to j's adjoint is gathered by j over its own list;
// every thread writes only its own particle and nothing is scattered.
// Ghosts carry no list of their own: each step their real partners are
// collected into per-ghost lists, sorted, and the ghost adjoints are mapped
// to the real parents with the mirror signs of the ghost fill.
//
// Per-particle stages: the constitutive chain (rhoupd, stress_rot1, f3dm10,
// sueos, nodalav, hieupd, eqos, stress_up, calccritts) is the forward's own
// code (mcm_physics.h) instantiated with dual numbers.  Its local Jacobian is
// taken column block by column block and applied transposed, so the yield
// active set, the EOS branch and the pressure cut-off are those the forward
// took.
//
// Scope: plain SPH (disctype 0), material models 10 (any EOS) and 1, fixed h
// (h_opt 0), tcrit_opt 0 or 1, no penalty contact, XSPH, dynamic relaxation
// or base acceleration.  With tcrit_opt 1 the critical step h / (c + |v|)
// reads |v| of the previous velocity update, which equals |v^n| (the
// reflections only flip signs), so it enters as a function of v^n.
// Parameters, per material: strinput(1:3) and eosinput(2) for model 10 (G,
// sigma_y, Eh, S1 of eos4), E and nu for model 1; and the deck's initial
// velocities.
#pragma once

constexpr int ADJ_NPAR = 4;                   // strinput[0], strinput[1], strinput[2], eosinput[1]
constexpr int ADJ_NIN = 20 + ADJ_NPAR;        // rod 6, spin 3, sigma 6, rho, e, efps, |v|, dt, parameters
constexpr int ADJ_K = 4;                      // directions per dual pass
enum { IN_ROD = 0, IN_SPIN = 6, IN_SIG = 9, IN_RHO = 15, IN_E = 16, IN_EFPS = 17, IN_VABS = 18, IN_DT = 19,
       IN_PAR = 20 };

// snapshot fields (np doubles each): everything a step reads of the real
// particles' state
enum {
    SN_X = 0, SN_V = 3, SN_R = 6, SN_H = 8, SN_HOLD = 10, SN_SIG = 11, SN_E = 17, SN_EFPS = 18, SN_P = 19,
    SN_C = 20, SN_Q = 21, SN_VABS = 22, SN_PCUT = 23, SN_PCUT2 = 24, SN_FAIL = 25, SN_DEFPS = 26,
    SN_ALFA = 27, SN_SD = 33, SN_N = 39
};

// adjoint scalars
enum { AS_DTB = 0, AS_DTP = 1, AS_CRIT = 2, AS_DTACC = 3, AS_ARGMIN = 4, AS_PAR = 8 };
constexpr int AS_N = AS_PAR + ADJ_NPAR * MCM_MAX_MAT;

// per-block partials: move (x.v, vbar.a, argmin), constitutive (dt, params), strain (dt)
enum { RA_XV = 0, RA_VA = 1, RA_ARG = 2, RA_CDT = 3, RA_PAR = 4, RA_SDT = RA_PAR + ADJ_NPAR * MCM_MAX_MAT };
constexpr int RA_N = RA_SDT + 1;

constexpr int ADJ_GOF = 8;                    // ghosts of one real particle (at most 7 for 3 planes)
// lanes per particle of the reverse pair gathers (a power of two up to 64):
// lane l of a particle's group takes list entries l, l + kLpp, ... and the
// group's partial sums are folded by a butterfly in a fixed order.  Eight
// lanes were 1.3 times as fast as a wavefront per particle on a
// 50,000-particle bar (the list's rows are 256 bytes apart, and a group of
// eight keeps eight particles' gathers in flight per wavefront).
constexpr int kLpp = 8;
__device__ __forceinline__ double group_sum(double v) {
    for (int off = kLpp / 2; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
    return v;
}

struct AdjDev {
    double *xb[3], *vb[3], *rb, *sb[6], *eb, *fb;   // adjoint of the state (real particles)
    double *abp, *qb, *rbn, *pdt;                   // adjoints local to a step; abp: acceleration [np][4]
    double* lbp;                                    // adjoint of the velocity gradient [np][10], g[n][m] at 3n + m
    double* xin;                                    // [ADJ_NIN][np] input adjoints of the constitutive reverse
    double *gxb[3], *gvb[3], *grb, *gAb[6];         // ghost adjoints [ngcap]
    int *gt_cnt, *gt, *gts;                         // real partners of each ghost [ngcap], [ngcap][gtcap],
                                                    // unsorted and sorted
    int *gof_cnt, *gof;                             // ghosts of each real particle [np], [np][ADJ_GOF]
    int gtcap, ngcap;
    const double* snap;                             // start-of-step snapshot of the step being reversed
    int cur;                                        // its h / rho buffer parity
    double* red;                                    // [nb][RA_N]
    double* as;                                     // [AS_N]
    int* err;                                       // bit 1: a transposed list overflowed
};

__device__ __forceinline__ double* snap_field(const DevState& d, int f) {
    if (f < SN_V) return d.x0[f];
    if (f < SN_R) return d.v[f - SN_V];
    if (f < SN_H) return d.rbuf[f - SN_R];
    if (f < SN_HOLD) return d.hbuf[f - SN_H];
    if (f == SN_HOLD) return d.hold;
    if (f < SN_E) return d.sig[f - SN_SIG];
    switch (f) {
        case SN_E: return d.e;
        case SN_EFPS: return d.efps;
        case SN_P: return d.p;
        case SN_C: return d.c;
        case SN_Q: return d.q;
        case SN_VABS: return d.vabs;
        case SN_PCUT: return d.pcut;
        case SN_PCUT2: return d.p_cut;
        case SN_FAIL: return d.fail;
        case SN_DEFPS: return d.defps_prev;
        default: break;
    }
    if (f < SN_SD) return d.alfa[f - SN_ALFA];
    return d.sd[f - SN_SD];
}

__device__ __forceinline__ double snap_at(const AdjDev& a, int np, int f, int i) { return a.snap[(long)f * np + i]; }

// packed per-particle adjoints read by the pair gathers (one aligned record
// per neighbour)
__device__ __forceinline__ void load_ab(const double* abp, int i, double v[3]) {
    const double2* p = reinterpret_cast<const double2*>(abp + 4L * i);
    const double2 x = p[0], y = p[1];
    v[0] = x.x; v[1] = x.y; v[2] = y.x;
}
__device__ __forceinline__ void load_g9(const double* lbp, int i, double G[9]) {
    const double2* p = reinterpret_cast<const double2*>(lbp + 10L * i);
    const double2 a = p[0], b = p[1], c = p[2], e = p[3], f = p[4];
    G[0] = a.x; G[1] = a.y; G[2] = b.x; G[3] = b.y; G[4] = c.x; G[5] = c.y; G[6] = e.x; G[7] = e.y; G[8] = f.x;
}

// save (1) or restore (0) the real-particle state; grid.y = field
// the state's fields (grid.y = field) and, in workgroup (0, 0), the
// scalars (ssc); nothing else here reads the scalars
__global__ __launch_bounds__(MCM_BLOCK) void k_snap(DevState d, double* base, Scalars* ssc, int save) {
    const int f = blockIdx.y;
    if (f == 0 && blockIdx.x == 0) {
        const int nw = (int)(sizeof(Scalars) / sizeof(int));
        int* dst = save ? reinterpret_cast<int*>(ssc) : reinterpret_cast<int*>(d.sc);
        const int* src = save ? reinterpret_cast<const int*>(d.sc) : reinterpret_cast<const int*>(ssc);
        for (int w = threadIdx.x; w < nw; w += MCM_BLOCK) dst[w] = src[w];
    }
    double* p = snap_field(d, f);
    double* b = base + (long)f * d.np;
    for (int i = blockIdx.x * MCM_BLOCK + threadIdx.x; i < d.np; i += gridDim.x * MCM_BLOCK) {
        if (save) b[i] = p[i];
        else p[i] = b[i];
    }
}

// positions of the sampled particles after a step
__global__ void k_adj_sample(DevState d, const int* sample, int ns, double* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const int i = sample[k];
    for (int c = 0; c < 3; c++) out[3 * k + c] = d.x0[c][i];
}

// xbar of the sampled particles += dL/dx of the recorded positions
__global__ void k_adj_seed(AdjDev a, const int* sample, int ns, const double* seed) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const int i = sample[k];
    for (int c = 0; c < 3; c++) a.xb[c][i] += seed[3 * k + c];
}

// ------------------------------------------------------------ kernel Hessian
// grad W of the cubic B-spline (mcm_gradw.f90) and its derivative for one
// pair, with two reciprocals (1/havg, 1/r): grad W = F xij with
// F = dw / (r havg), and H = d(grad W)/d(xij) = F I + C xij xij^T with
// C = (dw' - dw / r) / (r^2 havg), dw' = d(dw)/dr; zero at r = 0.  The reverse only
// needs the forward's gradient to rounding, not bit for bit.
struct KGH {
    double g[3], F, C;
};
__device__ __forceinline__ void kernel_gh(const double xij[3], double havg, double ipi, KGH& k) {
    const double r2 = xij[0] * xij[0] + xij[1] * xij[1] + xij[2] * xij[2];
    k.F = 0.0;
    k.C = 0.0;
    if (r2 > 0.0) {
        const double rad = sqrt(r2);
        const double ih = 1.0 / havg, ir = 1.0 / rad;
        const double z = rad * ih;
        const double c1 = ih * ih * ih * ipi;
        double dw = 0.0, dwz = 0.0;
        if (z < 1.0) {
            dw = c1 * ((2.25 * z - 3.0) * z);
            dwz = c1 * (4.5 * z - 3.0);
        } else if (z < 2.0) {
            const double t = 2.0 - z;
            dw = -(0.75 * c1 * (t * t));
            dwz = 1.5 * c1 * t;
        }
        k.F = dw * ir * ih;
        k.C = (dwz * ih - dw * ir) * ir * ir * ih;
    }
    for (int c = 0; c < 3; c++) k.g[c] = k.F * xij[c];
}
__device__ __forceinline__ void hess_apply(const KGH& k, const double xij[3], const double w[3], double out[3]) {
    const double xw = xij[0] * w[0] + xij[1] * w[1] + xij[2] * w[2];
    for (int c = 0; c < 3; c++) out[c] = k.F * w[c] + k.C * xw * xij[c];
}

// ------------------------------------------------------- ghost transposes
// Real partners of each ghost, from the real particles' lists, and the ghosts
// of each real particle, in one launch: blocks [0, nb) take the lists, the
// rest the ghosts.  A wavefront of the list part takes ADJ_GT_PPW particles
// and spreads their ghost entries (list slots nreal..ntot-1) over its 64
// lanes, ADJ_GT_U entries a lane per round, so that the loads and atomics of
// a round are in flight together.  Particles near a symmetry plane have tens
// of ghost entries each, and few particles per wavefront keep its rounds few.
// The order of the entries within a ghost's list is fixed by k_adj_sort.
constexpr int ADJ_GT_U = 8;
constexpr int ADJ_GT_PPW = 8;

__device__ __forceinline__ void gt_fill_wave(const DevState& d, const AdjDev& a, int wave) {
    const int lane = threadIdx.x & 63;
    const int wbase = wave * ADJ_GT_PPW, i = wbase + lane;
    int sp = 0, k0 = 0, n = 0;   // list position, first ghost slot, ghost entries
    if (lane < ADJ_GT_PPW && i < d.np) {
        sp = d.spos[i];
        k0 = d.nreal[i];
        n = d.ntot[i] - k0;
    }
    int inc = n;   // inclusive prefix of the entry counts over the wavefront
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    const int exc = inc - n;
    const int total = __shfl(inc, 63, 64);
    for (int t0 = 0; t0 < total; t0 += ADJ_GT_U * 64) {
        int g[ADJ_GT_U], who[ADJ_GT_U], slot[ADJ_GT_U];
#pragma unroll
        for (int u = 0; u < ADJ_GT_U; u++) {
            const int t = t0 + u * 64 + lane;
            // owner of entry t < total: the last lane whose exclusive prefix
            // is <= t (lane 0's is 0; lanes from ADJ_GT_PPW on hold total)
            int lo = 0;
            for (int step = ADJ_GT_PPW / 2; step > 0; step >>= 1)
                if (__shfl(exc, lo + step, 64) <= t) lo += step;
            const int k = __shfl(k0, lo, 64) + (t - __shfl(exc, lo, 64));
            const int row = __shfl(sp, lo, 64);
            who[u] = wbase + lo;
            g[u] = t < total ? nbr_row(d, row)[(long)k * 64] - d.np : -1;
        }
#pragma unroll
        for (int u = 0; u < ADJ_GT_U; u++) slot[u] = g[u] >= 0 ? atomicAdd(&a.gt_cnt[g[u]], 1) : 0;
#pragma unroll
        for (int u = 0; u < ADJ_GT_U; u++) {
            if (g[u] < 0) continue;
            if (slot[u] < a.gtcap) a.gt[(long)g[u] * a.gtcap + slot[u]] = who[u];
            else atomicOr(a.err, 1);
        }
    }
}

__global__ __launch_bounds__(MCM_BLOCK) void k_adj_fill(DevState d, AdjDev a, int nb) {
    if ((int)blockIdx.x < nb) {
        gt_fill_wave(d, a, blockIdx.x * (MCM_BLOCK / 64) + (threadIdx.x >> 6));
        return;
    }
    const int ngp = d.sc->ngp, gb = gridDim.x - nb;
    for (int g = (blockIdx.x - nb) * MCM_BLOCK + threadIdx.x; g < ngp; g += gb * MCM_BLOCK) {
        const int p = d.gpar[g];
        const int slot = atomicAdd(&a.gof_cnt[p], 1);
        if (slot < ADJ_GOF) a.gof[(long)p * ADJ_GOF + slot] = g;
        else atomicOr(a.err, 1);
    }
}

__device__ __forceinline__ void isort(int* v, int n, long stride) {
    for (int k = 1; k < n; k++) {
        const int x = v[k * stride];
        int j = k - 1;
        while (j >= 0 && v[j * stride] > x) {
            v[(j + 1) * stride] = v[j * stride];
            j--;
        }
        v[(j + 1) * stride] = x;
    }
}

// Both lists in ascending order, so that every sum over them has a fixed
// order, in one launch: blocks [0, wg) take the ghosts, one wavefront per
// ghost, each entry going to the slot given by the number of smaller ids in
// the list (the ids of one list are distinct); the rest take the real
// particles, one thread each.
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_sort(DevState d, AdjDev a, int wg) {
    if ((int)blockIdx.x >= wg) {
        const int i = (blockIdx.x - wg) * MCM_BLOCK + threadIdx.x;
        if (i < d.np) isort(a.gof + (long)i * ADJ_GOF, min(a.gof_cnt[i], ADJ_GOF), 1);
        return;
    }
    const int ngp = d.sc->ngp;
    const int lane = threadIdx.x & 63, wpb = MCM_BLOCK / 64;
    for (int g = blockIdx.x * wpb + (threadIdx.x >> 6); g < ngp; g += wg * wpb) {
        const int* src = a.gt + (long)g * a.gtcap;
        int* dst = a.gts + (long)g * a.gtcap;
        const int cnt = min(a.gt_cnt[g], a.gtcap);
        for (int k0 = 0; k0 < cnt; k0 += 64) {
            const int k = k0 + lane;
            const int v = k < cnt ? src[k] : 0;
            int rank = 0;
            for (int j = 0; j < cnt; j++) rank += src[j] < v ? 1 : 0;
            if (k < cnt) dst[rank] = v;
        }
    }
}

// ------------------------------------------------- move and velocity update
// Reverse of mcm_update_velocity + mcm_move + mcm_check_sym_pen for real
// particle i: v' = v + a dtn, x' = x1 + v' dt', then the reflections.
// xb, vb hold the adjoint of x', v' on entry and of x^n, v^n (partial) on
// exit; ab gets the adjoint of the acceleration after mcm_boundcond.
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_move(DevState d, AdjDev a) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    const Scalars* sc = d.sc;
    double pxv = 0.0, pva = 0.0, parg = 1.0e300;
    if (i < d.np) {
        const double dt1 = sc->dt, dt0 = sc->dtold;
        const double dtn = 0.5 * (dt1 + dt0);
        const SRec me = load_srec(d, i);
        const double vn[3] = {me.vx, me.vy, me.vz};
        double acc[3], vp[3], x[3], v[3], sx[3] = {1.0, 1.0, 1.0}, sv[3] = {1.0, 1.0, 1.0};
        const MRec mr = load_mrec(d, i);   // position after the strain pass (x1)
        const double x1[3] = {mr.x, mr.y, mr.z};
        for (int k = 0; k < 3; k++) {
            acc[k] = d.a[k][i];
            vp[k] = vn[k] + acc[k] * dtn;
            x[k] = x1[k] + vp[k] * dt1;
            v[k] = vp[k];
        }
        if (d.boundary) {   // the reflections of sym_pen, in its order
            for (int k = 0; k < 3; k++) {
                const double b0 = d.bx[k][0], b1 = d.bx[k][1];
                if (d.bcode[k][0] == 1 && x[k] < b0) { x[k] = 2.0 * b0 - x[k]; v[k] = -v[k]; sx[k] = -sx[k]; sv[k] = -sv[k]; }
                if (d.bcode[k][1] == 1 && x[k] > b1) { x[k] = 2.0 * b1 - x[k]; v[k] = -v[k]; sx[k] = -sx[k]; sv[k] = -sv[k]; }
                if (d.bcode[k][0] == 2) {
                    const double span = b1 - b0;
                    if (x[k] < b0) x[k] = x[k] + span;
                    if (x[k] > b1) x[k] = x[k] - span;
                }
            }
        }
        double abar[3];
        for (int k = 0; k < 3; k++) {
            const double xpb = sx[k] * a.xb[k][i];
            double vpb = sv[k] * a.vb[k][i] + dt1 * xpb;
            pxv += xpb * vp[k];
            pva += vpb * acc[k];
            abar[k] = dtn * vpb;
            a.xb[k][i] = xpb;
            a.vb[k][i] = vpb;
        }
        boundcond(d.dispbc[i], abar);
        double2* ap = reinterpret_cast<double2*>(a.abp + 4L * i);
        ap[0] = make_double2(abar[0], abar[1]);
        ap[1] = make_double2(abar[2], 0.0);
        if (d.critts_p[i] == sc->critts) parg = (double)i;
    }
    const double s1 = block_reduce<0>(pxv);
    const double s2 = block_reduce<0>(pva);
    const double s3 = block_reduce<1>(parg);
    if (threadIdx.x == 0) {
        double* r = a.red + (long)blockIdx.x * RA_N;
        r[RA_XV] = s1; r[RA_VA] = s2; r[RA_ARG] = s3;
    }
}

// dt' = min(tssf critts, 1.1 dt) (mcm_solution.f90:58-62) reversed, with the
// adjoint of dt' from the move and the velocity update.  One workgroup.
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_fold_move(DevState d, AdjDev a, int nblk) {
    const double xv = fold<0>(a.red, RA_N, RA_XV, nblk, 0.0);
    const double va = fold<0>(a.red, RA_N, RA_VA, nblk, 0.0);
    const double am = fold<1>(a.red, RA_N, RA_ARG, nblk, 1.0e300);
    if (threadIdx.x != 0) return;
    const Scalars* sc = d.sc;
    const double dtp = a.as[AS_DTB] + xv + 0.5 * va;
    double acc = 0.5 * va, crit = 0.0;
    if (d.tssf * sc->critts <= 1.1 * sc->dtold) crit = d.tssf * dtp;
    else acc = acc + 1.1 * dtp;
    a.as[AS_DTP] = dtp;
    a.as[AS_CRIT] = crit;
    a.as[AS_DTACC] = acc;
    a.as[AS_ARGMIN] = am;
}

// ------------------------------------------------------------- momentum
// Reverse of mcm_oldacceleration.f90:60-89 for real particle t, with the
// terms of t in its neighbours' sums gathered over t's own list:
//   a_t = sum_u m_u (A_u + A_t) grad W(x_t - x_u),
//   Abar_t += (m_u abar_t - m_t abar_u) (x) grad W, xbar_t += H (A_u + A_t) (m_u abar_t - m_t abar_u)
// (abar_u = 0 for ghosts).  A = (sigma - q I) / rho^2 maps to sigma', q', rho'.
// kLpp lanes per particle: lane l takes list entries l, l + kLpp, ...; the
// sums are combined across the group's lanes in a fixed butterfly order.
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_mom(DevState d, AdjDev a) {
    const double ipi = 1.0 / d.pi;
    const int lane = threadIdx.x & (kLpp - 1);
    const int t = blockIdx.x * (MCM_BLOCK / kLpp) + threadIdx.x / kLpp;
    if (t >= d.np) return;   // the particle's whole group
    const int nxt = a.cur ^ 1;
    const MRec me = load_mrec(d, t);
    double at[3];
    load_ab(a.abp, t, at);
    double Ab[6] = {0, 0, 0, 0, 0, 0}, xbt[3] = {0, 0, 0};
    const int n = d.ntot[t];
    const int* lst = nbr_row(d, d.spos[t]);
    for (int k = lane; k < n; k += kLpp) {
        const int u = lst[(long)k * 64];
        const MRec r = load_mrec(d, u);
        const double havg = 0.5 * (me.h + r.h);
        const double xij[3] = {me.x - r.x, me.y - r.y, me.z - r.z};
        KGH kg;
        kernel_gh(xij, havg, ipi, kg);
        const double* g = kg.g;
        double w[3], au[3] = {0.0, 0.0, 0.0};
        if (u < d.np) load_ab(a.abp, u, au);
        for (int c = 0; c < 3; c++) w[c] = r.m * at[c] - me.m * au[c];
        for (int nn = 0; nn < 3; nn++)
            for (int mm = 0; mm < 3; mm++) Ab[sym_index(mm, nn)] += w[nn] * g[mm];
        double bw[3];
        for (int mm = 0; mm < 3; mm++) {
            bw[mm] = 0.0;
            for (int nn = 0; nn < 3; nn++) {
                const int c = sym_index(mm, nn);
                bw[mm] += (r.A[c] + me.A[c]) * w[nn];
            }
        }
        double hw[3];
        hess_apply(kg, xij, bw, hw);
        for (int c = 0; c < 3; c++) xbt[c] += hw[c];
    }
    for (int c = 0; c < 6; c++) Ab[c] = group_sum(Ab[c]);
    for (int c = 0; c < 3; c++) xbt[c] = group_sum(xbt[c]);
    if (lane != 0) return;
    const double rho = d.rbuf[nxt][t], r2 = rho * rho;
    double rbar = 0.0;
    for (int c = 0; c < 6; c++) {
        a.sb[c][t] += Ab[c] / r2;
        rbar += Ab[c] * me.A[c];
    }
    a.qb[t] = -(Ab[0] + Ab[1] + Ab[2]) / r2;
    a.rb[t] += -2.0 * rbar / rho;
    for (int c = 0; c < 3; c++) a.xb[c][t] += xbt[c];
}

// the ghosts' share of the momentum reverse, gathered over their real partners
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_mom_ghost(DevState d, AdjDev a) {
    const double ipi = 1.0 / d.pi;
    const int ngp = d.sc->ngp;
    const int lane = threadIdx.x & 63, wpb = MCM_BLOCK / 64;
    for (int g = blockIdx.x * wpb + (threadIdx.x >> 6); g < ngp; g += gridDim.x * wpb) {
        const MRec me = load_mrec(d, d.np + g);
        double Ab[6] = {0, 0, 0, 0, 0, 0}, xbg[3] = {0, 0, 0};
        const int cnt = min(a.gt_cnt[g], a.gtcap);
        const int* lst = a.gts + (long)g * a.gtcap;
        for (int k = lane; k < cnt; k += 64) {
            const int i = lst[k];
            const MRec r = load_mrec(d, i);
            const double havg = 0.5 * (r.h + me.h);
            const double xij[3] = {r.x - me.x, r.y - me.y, r.z - me.z};
            KGH kg;
            kernel_gh(xij, havg, ipi, kg);
            const double* gw = kg.g;
            double w[3], ai[3];
            load_ab(a.abp, i, ai);
            for (int c = 0; c < 3; c++) w[c] = me.m * ai[c];
            for (int nn = 0; nn < 3; nn++)
                for (int mm = 0; mm < 3; mm++) Ab[sym_index(mm, nn)] += w[nn] * gw[mm];
            double bw[3];
            for (int mm = 0; mm < 3; mm++) {
                bw[mm] = 0.0;
                for (int nn = 0; nn < 3; nn++) {
                    const int c = sym_index(mm, nn);
                    bw[mm] += (me.A[c] + r.A[c]) * w[nn];
                }
            }
            double hw[3];
            hess_apply(kg, xij, bw, hw);
            for (int c = 0; c < 3; c++) xbg[c] -= hw[c];
        }
        for (int c = 0; c < 6; c++) Ab[c] = wave_sum(Ab[c]);
        for (int c = 0; c < 3; c++) xbg[c] = wave_sum(xbg[c]);
        if (lane == 0) {
            for (int c = 0; c < 6; c++) a.gAb[c][g] = Ab[c];
            for (int c = 0; c < 3; c++) a.gxb[c][g] = xbg[c];
        }
    }
}

// ghost momentum adjoints to the real parents: a ghost's A is its parent's
// (sigma' - q') with the off-diagonals negated, over the ghost's own
// start-of-step rho (mcm_update_ghost_stress.f90:28-41); its position is the
// parent's x^n mirrored
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_gmap_mom(DevState d, AdjDev a) {
    const int p = blockIdx.x * MCM_BLOCK + threadIdx.x;
    if (p >= d.np) return;
    const int cnt = min(a.gof_cnt[p], ADJ_GOF);
    for (int k = 0; k < cnt; k++) {
        const int g = a.gof[(long)p * ADJ_GOF + k], u = d.np + g;
        const MRec me = load_mrec(d, u);
        const double rg = d.rbuf[a.cur][u], r2 = rg * rg;
        double rbar = 0.0;
        for (int c = 0; c < 6; c++) {
            const double ab = a.gAb[c][g];
            a.sb[c][p] += (c < 3 ? ab : -ab) / r2;
            rbar += ab * me.A[c];
        }
        a.qb[p] += -(a.gAb[0][g] + a.gAb[1][g] + a.gAb[2][g]) / r2;
        a.rbn[p] += -2.0 * rbar / rg;
        const int sg = d.gsgn[g];
        for (int c = 0; c < 3; c++) a.xb[c][p] += ((sg >> c) & 1 ? -a.gxb[c][g] : a.gxb[c][g]);
    }
}

// ---------------------------------------------------------- constitutive
// The material constants the chains read (f3dm10: strinput(1:8); f3dm1:
// strinput(1), (6), (20); the EOS routines: eosinput(1:8); rhoupd, nodalav,
// sueos: rho, h, rho_max, av_l, av_q, g), as duals.
template <int K>
struct MatDual {
    int model, eos, visc_type, ispall, efps_flag;
    Dual<K> rho, g, av_l, av_q, h, rho_max;
    Dual<K> strinput[20], eosinput[8];
};

template <int K>
__device__ __forceinline__ MatDual<K> dual_mat(const MatParams& m) {
    MatDual<K> r;
    r.model = m.model; r.eos = m.eos; r.visc_type = m.visc_type; r.ispall = m.ispall; r.efps_flag = m.efps_flag;
    r.rho = m.rho; r.g = m.g; r.av_l = m.av_l; r.av_q = m.av_q; r.h = m.h; r.rho_max = m.rho_max;
#pragma unroll
    for (int k = 0; k < 20; k++) r.strinput[k] = m.strinput[k];
#pragma unroll
    for (int k = 0; k < 8; k++) r.eosinput[k] = m.eosinput[k];
    return r;
}

template <int K>
__device__ __forceinline__ void seed(Dual<K>& x, int j, int p0) {
    if (j >= p0 && j < p0 + K) x.d[j - p0] = 1.0;
}

constexpr int ADJ_NPASS = (ADJ_NIN + ADJ_K - 1) / ADJ_K;

// model 1: the set-up constants from E and nu (mcm_sets1.f90:23-51, as
// setup_material derives them): G = q2 at strinput(20) and lambda + 2G = q3
// at strinput(2)
template <class T, class M>
MCM_HD void model1_constants(M& m, const T& E, const T& nu) {
    const T q1 = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const T q2 = E * 0.5 / (1.0 + nu);
    const T q3 = q1 + 2.0 * q2;
    m.strinput[0] = E;
    m.strinput[5] = nu;
    m.strinput[19] = q2;
    m.strinput[1] = q3;
}

// One dual pass of the reverse of the constitutive update of real particle i
// (mcm_constitutive.f90 with mcm_rhoupd, mcm_stress_rot1, mcm_f3dm10,
// mcm_sueos, mcm_nodalav, mcm_hieupd, mcm_eqos, mcm_stress_up and
// mcm_calccritts): the forward's chain with the inputs P0 .. P0 + K - 1 seeded
// gives K columns of the local Jacobian of (sigma', rho', e', efps', q',
// critts) with respect to (rod, spin, sigma, rho, e, efps, |v|, dt,
// parameters); out[k] = adjoint of input P0 + k.
template <int P0>
__device__ __forceinline__ void const_pass(const DevState& d, const AdjDev& a, int i, double out[ADJ_K]) {
    typedef Dual<ADJ_K> D;
    const int np = d.np;
    const MatParams& m0 = d.mat[d.pmat[i]];
    const double dtn = d.sc->dtold;   // the step's dt (before the dt update)
    PStateT<D> s;
    s.mass = d.mass[i];
    s.rho0 = d.rho0[i];
    s.h = snap_at(a, np, SN_H + a.cur, i);
    s.hold = snap_at(a, np, SN_HOLD, i);
    s.h0 = d.h0[i];
    s.p = snap_at(a, np, SN_P, i);
    s.c = snap_at(a, np, SN_C, i);
    s.vabs = snap_at(a, np, SN_VABS, i);
    seed(s.vabs, IN_VABS, P0);
    s.pcut = snap_at(a, np, SN_PCUT, i);
    s.p_cut = snap_at(a, np, SN_PCUT2, i);
    s.fail = snap_at(a, np, SN_FAIL, i);
    s.q = snap_at(a, np, SN_Q, i);
    s.einc = 0.0;
    s.etry = 0.0;
#pragma unroll
    for (int c = 0; c < 6; c++) { s.rod[c] = d.rod[c][i]; seed(s.rod[c], IN_ROD + c, P0); }
#pragma unroll
    for (int c = 0; c < 3; c++) { s.spin[c] = d.spin[c][i]; seed(s.spin[c], IN_SPIN + c, P0); }
#pragma unroll
    for (int c = 0; c < 6; c++) {
        s.sig[c] = snap_at(a, np, SN_SIG + c, i);
        seed(s.sig[c], IN_SIG + c, P0);
    }
    s.rho = snap_at(a, np, SN_R + a.cur, i);
    seed(s.rho, IN_RHO, P0);
    s.e = snap_at(a, np, SN_E, i);
    seed(s.e, IN_E, P0);
    s.efps = snap_at(a, np, SN_EFPS, i);
    seed(s.efps, IN_EFPS, P0);
    D dt = dtn;
    seed(dt, IN_DT, P0);
    MatDual<ADJ_K> m = dual_mat<ADJ_K>(m0);
    s.tracerod = s.rod[S11] + s.rod[S22] + s.rod[S33];
    if (m0.model == 10) {
        seed(m.strinput[0], IN_PAR + 0, P0);
        seed(m.strinput[1], IN_PAR + 1, P0);
        seed(m.strinput[2], IN_PAR + 2, P0);
        seed(m.eosinput[1], IN_PAR + 3, P0);
        m.g = m.strinput[0];   // mcm_calcg.f90: G of model 10
        // the forward's chain (constitutive_particle, model 10)
        rhoupd(s, m, dt);
        stress_rot1(s, dt);
        double dprev = snap_at(a, np, SN_DEFPS, i);
        f3dm10(s, m, dt, dprev);
        sueos(s, m);
        nodalav(s, m);
        hieupd(s, dt);
        eqos(s, m, dt);
        stress_up(s);
    } else {
        // model 1 (constitutive_particle, model 1): parameters E and nu; the
        // sound speed sqrt((lambda + 2G) / rho) is fixed at the initialisation
        // (mcm_calcrhopc.f90:51), so its derivative follows lambda + 2G
        D E = m0.strinput[0], nu = m0.strinput[5];
        seed(E, IN_PAR + 0, P0);
        seed(nu, IN_PAR + 1, P0);
        model1_constants(m, E, nu);
        const double c0 = s.c.v, q3 = m0.strinput[1];
#pragma unroll
        for (int k = 0; k < ADJ_K; k++) s.c.d[k] = c0 / (2.0 * q3) * m.strinput[1].d[k];
        rhoupd(s, m, dt);
        stress_rot1(s, dt);
        f3dm1(s, m, dt);
        nodalav(s, m);
        lieupd(s, dt);
    }
    const D cr = critts(s, m, d.tcrit_opt, 0.0);
    double ys[6];
#pragma unroll
    for (int c = 0; c < 6; c++) ys[c] = a.sb[c][i];
    const double yr = a.rb[i], ye = a.eb[i], yf = a.fb[i], yq = a.qb[i];
    const double yc = (double)i == a.as[AS_ARGMIN] ? a.as[AS_CRIT] : 0.0;
#pragma unroll
    for (int k = 0; k < ADJ_K; k++) {
        double v = yr * s.rho.d[k] + ye * s.e.d[k] + yf * s.efps.d[k] + yq * s.q.d[k] + yc * cr.d[k];
#pragma unroll
        for (int c = 0; c < 6; c++) v += ys[c] * s.sig[c].d[k];
        out[k] = v;
    }
}

// The passes of the constitutive reverse, one thread per (particle, pass):
// grid.y = pass; the input adjoints go to the staging array xin[input][np].
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_const(DevState d, AdjDev a) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    if (i >= d.np) return;
    double out[ADJ_K];
    switch (blockIdx.y) {
        case 0: const_pass<0 * ADJ_K>(d, a, i, out); break;
        case 1: const_pass<1 * ADJ_K>(d, a, i, out); break;
        case 2: const_pass<2 * ADJ_K>(d, a, i, out); break;
        case 3: const_pass<3 * ADJ_K>(d, a, i, out); break;
        case 4: const_pass<4 * ADJ_K>(d, a, i, out); break;
        default: const_pass<5 * ADJ_K>(d, a, i, out); break;
    }
    static_assert(ADJ_NPASS == 6, "one case per pass");
#pragma unroll
    for (int k = 0; k < ADJ_K; k++) {
        const int j = blockIdx.y * ADJ_K + k;
        if (j < ADJ_NIN) a.xin[(long)j * d.np + i] = out[k];
    }
}

// Input adjoints of the constitutive reverse to the state adjoints: rod /
// spin -> velocity gradient g (rod_spin: rod_ab = (g_ab + g_ba) / 2,
// spin = (g_ab - g_ba) / 2 for (ab) = (01), (02), (12)), |v^n| -> v^n
// (mcm_update_velocity.f90:45-50), per-block partials of dt and the
// parameters.
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_const_fin(DevState d, AdjDev a) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    const int np = d.np;
    const int mi = i < np ? d.pmat[i] : 0;
    double x[ADJ_NIN];
#pragma unroll
    for (int j = 0; j < ADJ_NIN; j++) x[j] = i < np ? a.xin[(long)j * np + i] : 0.0;
    if (i < np) {
        const double* rb_ = x + IN_ROD;
        const double* wb = x + IN_SPIN;
        double G[10];
        G[0] = rb_[S11];
        G[4] = rb_[S22];
        G[8] = rb_[S33];
        G[1] = 0.5 * rb_[S12] + 0.5 * wb[0];
        G[3] = 0.5 * rb_[S12] - 0.5 * wb[0];
        G[2] = 0.5 * rb_[S13] + 0.5 * wb[1];
        G[6] = 0.5 * rb_[S13] - 0.5 * wb[1];
        G[5] = 0.5 * rb_[S23] + 0.5 * wb[2];
        G[7] = 0.5 * rb_[S23] - 0.5 * wb[2];
        G[9] = 0.0;
        double2* gp = reinterpret_cast<double2*>(a.lbp + 10L * i);
        for (int c = 0; c < 5; c++) gp[c] = make_double2(G[2 * c], G[2 * c + 1]);
        for (int c = 0; c < 6; c++) a.sb[c][i] = x[IN_SIG + c];
        a.rb[i] = x[IN_RHO];
        a.eb[i] = x[IN_E];
        a.fb[i] = x[IN_EFPS];
        const double va = snap_at(a, np, SN_VABS, i);
        if (x[IN_VABS] != 0.0 && va > 0.0) {
            const SRec me = load_srec(d, i);
            const double f = x[IN_VABS] / va;
            a.vb[0][i] += f * me.vx;
            a.vb[1][i] += f * me.vy;
            a.vb[2][i] += f * me.vz;
        }
    }
    const double sdt = block_reduce<0>(x[IN_DT]);
    double* r = a.red + (long)blockIdx.x * RA_N;
    if (threadIdx.x == 0) r[RA_CDT] = sdt;
    for (int m = 0; m < d.nummat; m++)
        for (int k = 0; k < ADJ_NPAR; k++) {
            const double s = block_reduce<0>(i < np && mi == m ? x[IN_PAR + k] : 0.0);
            if (threadIdx.x == 0) r[RA_PAR + ADJ_NPAR * m + k] = s;
        }
}

// --------------------------------------------------------------- strain
// Reverse of mcm_oldstrain.f90:34-64 with the half-step positions of
// mcm_strain.f90:27-31 for real particle t, with t's terms in its real
// neighbours' sums gathered over t's own list:
//   g_t[n][m] = sum_u V_u (v_u - v_t)_n grad_m W(xh_t - xh_u), V = m / rho.
// kLpp lanes per particle, as k_adj_mom; the particle's share of the dt
// adjoint goes to pdt (reduced by k_adj_gmap_strain).
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_strain(DevState d, AdjDev a) {
    const double ipi = 1.0 / d.pi;
    const int lane = threadIdx.x & (kLpp - 1);
    const int t = blockIdx.x * (MCM_BLOCK / kLpp) + threadIdx.x / kLpp;
    if (t >= d.np) return;   // the particle's whole group
    const double dt = d.sc->dtold;
    const SRec me = load_srec(d, t);   // records hold the strain pass's positions
    const double vt[3] = {me.vx, me.vy, me.vz};
    const double xh[3] = {me.x, me.y, me.z};
    const double rho_t = d.rbuf[a.cur][t];
    double Gt[9];
    load_g9(a.lbp, t, Gt);
    double vbt[3] = {0, 0, 0}, xhb[3] = {0, 0, 0}, rbt = 0.0;
    const int n = d.ntot[t];
    const int* lst = nbr_row(d, d.spos[t]);
    for (int k = lane; k < n; k += kLpp) {
        const int u = lst[(long)k * 64];
        const SRec r = load_srec(d, u);
        const double xj[3] = {r.x, r.y, r.z};
        const double havg = 0.5 * (me.h + r.h);
        const double xij[3] = {xh[0] - xj[0], xh[1] - xj[1], xh[2] - xj[2]};
        KGH kg;
        kernel_gh(xij, havg, ipi, kg);
        const double* g = kg.g;
        const double dv[3] = {r.vx - vt[0], r.vy - vt[1], r.vz - vt[2]};
        // own terms
        double w[3];
        for (int mm = 0; mm < 3; mm++) w[mm] = Gt[mm] * dv[0] + Gt[3 + mm] * dv[1] + Gt[6 + mm] * dv[2];
        for (int nn = 0; nn < 3; nn++)
            vbt[nn] -= r.vol * (Gt[3 * nn] * g[0] + Gt[3 * nn + 1] * g[1] + Gt[3 * nn + 2] * g[2]);
        double hw[3];
        hess_apply(kg, xij, w, hw);
        for (int c = 0; c < 3; c++) xhb[c] += r.vol * hw[c];
        // t's terms in the sum of real neighbour u
        if (u < d.np) {
            double Gu[9];
            load_g9(a.lbp, u, Gu);
            double w2[3];
            for (int mm = 0; mm < 3; mm++) w2[mm] = -(Gu[mm] * dv[0] + Gu[3 + mm] * dv[1] + Gu[6 + mm] * dv[2]);
            for (int nn = 0; nn < 3; nn++)
                vbt[nn] -= me.vol * (Gu[3 * nn] * g[0] + Gu[3 * nn + 1] * g[1] + Gu[3 * nn + 2] * g[2]);
            rbt += me.vol / rho_t * (w2[0] * g[0] + w2[1] * g[1] + w2[2] * g[2]);
            hess_apply(kg, xij, w2, hw);
            for (int c = 0; c < 3; c++) xhb[c] -= me.vol * hw[c];
        }
    }
    for (int c = 0; c < 3; c++) {
        vbt[c] = group_sum(vbt[c]);
        xhb[c] = group_sum(xhb[c]);
    }
    rbt = group_sum(rbt);
    if (lane != 0) return;
    // xh = x - v dt / 2
    double pdt = 0.0;
    for (int c = 0; c < 3; c++) {
        a.xb[c][t] += xhb[c];
        a.vb[c][t] += vbt[c] - 0.5 * dt * xhb[c];
        pdt += -0.5 * vt[c] * xhb[c];
    }
    a.rbn[t] += rbt;
    a.pdt[t] = pdt;
}

// the ghosts' share of the strain reverse (ghost positions are not shifted)
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_strain_ghost(DevState d, AdjDev a) {
    const double ipi = 1.0 / d.pi;
    const int ngp = d.sc->ngp;
    const int lane = threadIdx.x & 63, wpb = MCM_BLOCK / 64;
    for (int g = blockIdx.x * wpb + (threadIdx.x >> 6); g < ngp; g += gridDim.x * wpb) {
        const int u = d.np + g;
        const SRec me = load_srec(d, u);
        const double rho_g = d.rbuf[a.cur][u];
        double vbg[3] = {0, 0, 0}, xbg[3] = {0, 0, 0}, rbg = 0.0;
        const int cnt = min(a.gt_cnt[g], a.gtcap);
        const int* lst = a.gts + (long)g * a.gtcap;
        for (int k = lane; k < cnt; k += 64) {
            const int i = lst[k];
            const SRec r = load_srec(d, i);
            const double xh[3] = {r.x, r.y, r.z};
            const double havg = 0.5 * (r.h + me.h);
            const double xij[3] = {xh[0] - me.x, xh[1] - me.y, xh[2] - me.z};
            KGH kg;
            kernel_gh(xij, havg, ipi, kg);
            const double* gw = kg.g;
            double Gi[9];
            load_g9(a.lbp, i, Gi);
            const double dv[3] = {me.vx - r.vx, me.vy - r.vy, me.vz - r.vz};
            double w[3];
            for (int mm = 0; mm < 3; mm++) w[mm] = Gi[mm] * dv[0] + Gi[3 + mm] * dv[1] + Gi[6 + mm] * dv[2];
            for (int nn = 0; nn < 3; nn++)
                vbg[nn] += me.vol * (Gi[3 * nn] * gw[0] + Gi[3 * nn + 1] * gw[1] + Gi[3 * nn + 2] * gw[2]);
            rbg += -(me.vol / rho_g) * (w[0] * gw[0] + w[1] * gw[1] + w[2] * gw[2]);
            double hw[3];
            hess_apply(kg, xij, w, hw);
            for (int c = 0; c < 3; c++) xbg[c] -= me.vol * hw[c];
        }
        for (int c = 0; c < 3; c++) {
            vbg[c] = wave_sum(vbg[c]);
            xbg[c] = wave_sum(xbg[c]);
        }
        rbg = wave_sum(rbg);
        if (lane == 0) {
            for (int c = 0; c < 3; c++) { a.gvb[c][g] = vbg[c]; a.gxb[c][g] = xbg[c]; }
            a.grb[g] = rbg;
        }
    }
}

// ghost strain adjoints to the real parents: x, v mirrored, rho copied
// (mcm_ghost_setup.f90:330-341)
// (and the block partials of the strain pass's dt adjoint)
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_gmap_strain(DevState d, AdjDev a) {
    const int p = blockIdx.x * MCM_BLOCK + threadIdx.x;
    double pdt = 0.0;
    if (p < d.np) {
        const int cnt = min(a.gof_cnt[p], ADJ_GOF);
        for (int k = 0; k < cnt; k++) {
            const int g = a.gof[(long)p * ADJ_GOF + k];
            const int sg = d.gsgn[g];
            for (int c = 0; c < 3; c++) {
                const double s = (sg >> c) & 1 ? -1.0 : 1.0;
                a.xb[c][p] += s * a.gxb[c][g];
                a.vb[c][p] += s * a.gvb[c][g];
            }
            a.rbn[p] += a.grb[g];
        }
        a.rb[p] += a.rbn[p];
        a.rbn[p] = 0.0;
        pdt = a.pdt[p];
    }
    const double sdt = block_reduce<0>(pdt);
    if (threadIdx.x == 0) a.red[(long)blockIdx.x * RA_N + RA_SDT] = sdt;
}

// the transposed lists' counts to zero before a reverse step fills them
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_zero_counts(AdjDev a, long ng, long np) {
    for (long k = blockIdx.x * (long)MCM_BLOCK + threadIdx.x; k < std::max(ng, np); k += (long)gridDim.x * MCM_BLOCK) {
        if (k < ng) a.gt_cnt[k] = 0;
        if (k < np) a.gof_cnt[k] = 0;
    }
}

// a recomputed step that stopped on a capacity (cannot happen: the forward ran
// the same states at the same or smaller capacities) is reported at the end
__global__ void k_adj_check(DevState d, AdjDev a) {
    if (d.sc->abort) atomicOr(a.err, 2);
}

// adjoint of the step's dt and of the parameters.  One workgroup.
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_fold_end(DevState d, AdjDev a, int nblk) {
    const double cdt = fold<0>(a.red, RA_N, RA_CDT, nblk, 0.0);
    const double sdt = fold<0>(a.red, RA_N, RA_SDT, nblk, 0.0);
    double par[ADJ_NPAR * MCM_MAX_MAT];
    for (int m = 0; m < d.nummat; m++)
        for (int k = 0; k < ADJ_NPAR; k++)
            par[ADJ_NPAR * m + k] = fold<0>(a.red, RA_N, RA_PAR + ADJ_NPAR * m + k, nblk, 0.0);
    if (threadIdx.x != 0) return;
    a.as[AS_DTB] = a.as[AS_DTACC] + cdt + sdt;
    for (int m = 0; m < d.nummat; m++)
        for (int k = 0; k < ADJ_NPAR; k++) a.as[AS_PAR + ADJ_NPAR * m + k] += par[ADJ_NPAR * m + k];
}


// The intermediates of one recomputed step that the adjoint kernels read:
// gather records, both density buffers (over reals and the step's ghosts),
// rod, spin, acceleration and critts of the reals; neighbour lists and
// counts, ghost parents and mirror signs, the lists' positions (spos).
// Field f of a slot: doubles at
// dbase + off, ints at ibase + off.
struct InterLayout {
    long ucap, np, nbrlen, ngcap;
};
constexpr int INTER_ND = 17, INTER_NI = 6;
// doubles per unified entry: the two records, then both density buffers
constexpr long INTER_RS = 2 * RS_SLOTS, INTER_RM = 2 * RM_SLOTS, INTER_UCAP = INTER_RS + INTER_RM + 2;

__device__ __forceinline__ void inter_field(const DevState& d, const InterLayout& L, int f, int ngp, double*& dp,
                                            int*& ip, long& len, long& off) {
    const long u = L.np + ngp;
    dp = nullptr;
    ip = nullptr;
    switch (f) {
        case 0: dp = d.rs; len = INTER_RS * u; off = 0; return;
        case 1: dp = d.rm; len = INTER_RM * u; off = INTER_RS * L.ucap; return;
        case 2: dp = d.rbuf[0]; len = u; off = (INTER_RS + INTER_RM) * L.ucap; return;
        case 3: dp = d.rbuf[1]; len = u; off = (INTER_RS + INTER_RM + 1) * L.ucap; return;
        default: break;
    }
    if (f < INTER_ND) {
        const int k = f - 4;   // rod 6, spin 3, a 3, critts
        dp = k < 6 ? d.rod[k] : (k < 9 ? d.spin[k - 6] : (k < 12 ? d.a[k - 9] : d.critts_p));
        len = L.np;
        off = INTER_UCAP * L.ucap + (long)k * L.np;
        return;
    }
    switch (f - INTER_ND) {
        case 0: ip = d.nbr; len = L.nbrlen; off = 0; return;   // rows below max_count only (k_inter)
        case 1: ip = d.nreal; len = L.np; off = L.nbrlen; return;
        case 2: ip = d.ntot; len = L.np; off = L.nbrlen + L.np; return;
        case 3: ip = d.gpar; len = ngp; off = L.nbrlen + 2 * L.np; return;
        case 4: ip = d.gsgn; len = ngp; off = L.nbrlen + 2 * L.np + L.ngcap; return;
        default: ip = d.spos; len = L.np; off = L.nbrlen + 2 * L.np + 2 * L.ngcap; return;   // list positions
    }
}

// save (1) or restore (0) the intermediates (grid.y = field) and the step's
// scalars after the step (ssc, copied by workgroup (0, 0)); the counts
// (ghosts, rows) come from the live scalars when saving and from the slot's
// when restoring
__global__ __launch_bounds__(MCM_BLOCK) void k_inter(DevState d, InterLayout L, double* dbase, int* ibase,
                                                     Scalars* ssc, int save, int skip_rows) {
    const Scalars& S = save ? *d.sc : *ssc;
    if (blockIdx.x == 0 && blockIdx.y == 0) {
        const int nw = (int)(sizeof(Scalars) / sizeof(int));
        int* dst = save ? reinterpret_cast<int*>(ssc) : reinterpret_cast<int*>(d.sc);
        const int* src = save ? reinterpret_cast<const int*>(d.sc) : reinterpret_cast<const int*>(ssc);
        for (int w = threadIdx.x; w < nw; w += MCM_BLOCK) dst[w] = src[w];
    }
    double* dp;
    int* ip;
    long len, off;
    inter_field(d, L, blockIdx.y, S.ngp, dp, ip, len, off);
    if (blockIdx.y == INTER_ND) {
        if (skip_rows) return;   // the rows stay in the search slot of the step
        // the blocked lists [npad/64][ncap][64]: rows k < the step's largest
        // count, a contiguous run of each 64-particle block
        const long rows = min(S.max_count, d.ncap), nblk64 = L.nbrlen / ((long)d.ncap * 64);
        for (long b = blockIdx.x; b < nblk64; b += gridDim.x) {
            const long q0 = b * d.ncap * 64;
            for (long k = threadIdx.x; k < rows * 64; k += MCM_BLOCK) {
                if (save) ibase[off + q0 + k] = ip[q0 + k];
                else ip[q0 + k] = ibase[off + q0 + k];
            }
        }
        return;
    }
    for (long i = blockIdx.x * (long)MCM_BLOCK + threadIdx.x; i < len; i += (long)gridDim.x * MCM_BLOCK) {
        if (dp) {
            if (save) dbase[off + i] = dp[i];
            else dp[i] = dbase[off + i];
        } else {
            if (save) ibase[off + i] = ip[i];
            else ip[i] = ibase[off + i];
        }
    }
}

// One recorded step's neighbour-search outputs, everything of the search
// that later kernels read: the blocked lists' rows below the step's longest
// list, nreal, ntot, mindist, the thread order (sorted, spos), the real
// particles' strain records, the ghosts (position, half-step position,
// velocity, mass, both h and rho buffers, material, parent, mirror signs,
// strain records and momentum records) and the range flags; the scalars
// after the search are
// kept beside it.  Slots are sized for the run's capacities (npad, ncap,
// ngcap) when the recording starts and used only while they are unchanged.
// The list rows are not copied: while a recorded step runs, and while it is
// recomputed and reversed, the run's list array is the slot's rows, which
// the list build writes and the pair and reverse kernels read in place.
struct SearchSlot {
    int *rows, *nreal, *ntot, *sorted, *spos, *gpmat, *gpar, *gsgn, *range;
    double *mindist, *rs, *rmg, *gd;   // rmg: the ghosts' momentum records; gd: [14][ngcap] x0 3, x1 3, v 3,
                                       // mass, hbuf 2, rbuf 2
};
constexpr int SNAP_GD = 14;

__device__ __forceinline__ double* snap_gfield(const DevState& d, int f) {
    if (f < 3) return d.x0[f];
    if (f < 6) return d.x1[f - 3];
    if (f < 9) return d.v[f - 6];
    if (f == 9) return d.mass;
    if (f < 12) return d.hbuf[f - 10];
    return d.rbuf[f - 12];
}

// save (1) or restore (0) one step's search outputs but the list rows (in
// place, above) and the scalars after the search (ssc, copied by workgroup
// 0).  The ghost count comes from the live scalars when saving and from the
// slot's when restoring, which workgroup 0 writes back meanwhile.
__global__ __launch_bounds__(MCM_BLOCK) void k_adj_search(DevState d, SearchSlot p, Scalars* ssc, int save) {
    const Scalars& S = save ? *d.sc : *ssc;
    const long np = d.np, ngp = S.ngp, gcap = d.ngcap;
    if (blockIdx.x == 0) {
        const int nw = (int)(sizeof(Scalars) / sizeof(int));
        int* dst = save ? reinterpret_cast<int*>(ssc) : reinterpret_cast<int*>(d.sc);
        const int* src = save ? reinterpret_cast<const int*>(d.sc) : reinterpret_cast<const int*>(ssc);
        for (int w = threadIdx.x; w < nw; w += MCM_BLOCK) dst[w] = src[w];
    }
    const long tid = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long nt = (long)gridDim.x * blockDim.x;
    for (long i = tid; i < np; i += nt) {
        if (save) {
            p.nreal[i] = d.nreal[i]; p.ntot[i] = d.ntot[i]; p.sorted[i] = d.sorted[i]; p.spos[i] = d.spos[i];
            p.mindist[i] = d.mindist[i];
        } else {
            d.nreal[i] = p.nreal[i]; d.ntot[i] = p.ntot[i]; d.sorted[i] = p.sorted[i]; d.spos[i] = p.spos[i];
            d.mindist[i] = p.mindist[i];
        }
    }
    // strain records of the reals and the step's ghosts, 16 bytes at a time
    const long nrs = RS_SLOTS * (np + ngp);
    for (long k = tid; k < nrs; k += nt) {
        if (save) reinterpret_cast<double2*>(p.rs)[k] = reinterpret_cast<const double2*>(d.rs)[k];
        else reinterpret_cast<double2*>(d.rs)[k] = reinterpret_cast<const double2*>(p.rs)[k];
    }
    const long nrm = RM_SLOTS * ngp, rm0 = RM_SLOTS * np;
    for (long k = tid; k < nrm; k += nt) {
        if (save) reinterpret_cast<double2*>(p.rmg)[k] = reinterpret_cast<const double2*>(d.rm)[rm0 + k];
        else reinterpret_cast<double2*>(d.rm)[rm0 + k] = reinterpret_cast<const double2*>(p.rmg)[k];
    }
    for (long g = tid; g < ngp; g += nt) {
        const long u = np + g;
        if (save) {
            p.gpmat[g] = d.pmat[u]; p.gpar[g] = d.gpar[g]; p.gsgn[g] = d.gsgn[g];
        } else {
            d.pmat[u] = p.gpmat[g]; d.gpar[g] = p.gpar[g]; d.gsgn[g] = p.gsgn[g];
        }
        for (int f = 0; f < SNAP_GD; f++) {
            double* a = snap_gfield(d, f);
            if (save) p.gd[f * gcap + g] = a[u];
            else a[u] = p.gd[f * gcap + g];
        }
    }
    if (tid == 0) {
        if (save) p.range[0] = d.range[0];
        else d.range[0] = p.range[0];
    }
}

// ================================================================== host
// Per-kernel timing of the reverse step, for profiling.  With MCM_ADJ_TIMING=1
// in the environment when an adjoint is created, an event is recorded before
// the first and after every kernel of each of its reverse steps; the
// intervals (each one kernel and the launch gap before it) are summed per
// kernel and printed to stderr when the adjoint is destroyed.  The events are
// read at the end of every reverse step, which holds the host there.
constexpr int ADJ_TK = 14;
const char* const adj_tk_names[ADJ_TK] = {
    "k_adj_zero_counts", "k_adj_fill", "k_adj_sort",  "k_adj_move",         "k_adj_fold_move",
    "k_adj_mom",         "k_adj_mom_ghost", "k_adj_gmap_mom", "k_adj_const", "k_adj_const_fin",
    "k_adj_strain",      "k_adj_strain_ghost", "k_adj_gmap_strain", "k_adj_fold_end"};
struct AdjTiming {
    hipEvent_t ev[ADJ_TK + 1] = {};
    double ms[ADJ_TK] = {};
    long steps = 0;
};

struct mcm_gpu_adjoint {
    mcm_gpu_ctx* ctx = nullptr;
    int nsteps = 0, k = 1, nseg = 0, nsample = 0;
    bool store = false;               // k == 1: every step's state and intermediates kept from the recording
    int ninter = 0;                   // intermediate slots (k, or nsteps when store)
    long np = 0;
    double* ckpt = nullptr;           // [nseg][SN_N][np]: state at the start of each segment
    Scalars* ckpt_sc = nullptr;       // [nseg]
    std::vector<int> ckpt_cur;
    double* seg = nullptr;            // [k][SN_N][np]: start of every step of the segment being reversed
                                      // (allocated by the first reverse sweep)
    Scalars* seg_sc = nullptr;
    std::vector<int> seg_cur;
    InterLayout lay{};                // intermediates of every step of the segment
    double* inter_d = nullptr;        // [k][INTER doubles]
    int* inter_i = nullptr;           // [k][INTER ints]
    Scalars* inter_sc = nullptr;      // [k]: scalars after the step (dt, dtold, critts, ngp)
    size_t inter_dn = 0, inter_in = 0;
    std::vector<void*> iallocs;       // intermediate slots (sized by the run's capacities)
    int* d_sample = nullptr;
    double* d_hist = nullptr;         // [nsteps][nsample][3]: positions, then their adjoints
    // every recorded step's neighbour-search outputs (k > 1, when they fit),
    // so that the reverse sweep's recomputation restores them instead of
    // searching again: the same lists, order, ghosts and records
    int* snap_i = nullptr;            // [nsteps][snap_ni]
    double* snap_d = nullptr;         // [nsteps][snap_nd]
    Scalars* snap_sc = nullptr;       // [nsteps]: the scalars after the search
    size_t snap_ni = 0, snap_nd = 0;
    int snap_npad = 0, snap_ncap = 0, snap_ngcap = 0;   // capacities the slots are sized for
    bool snap_ok = false;             // every step of the last recording saved at those capacities
    int* nbr_own = nullptr;           // the run's own list array while a slot's rows stand in for it
    std::vector<void*> lallocs;
    AdjDev a{};
    std::vector<void*> allocs, gallocs;   // run-long and ghost-sized allocations
    AdjTiming* timing = nullptr;          // MCM_ADJ_TIMING=1 only
    std::string err;
};

namespace {

#define HIPCHK_A(call)                                                                            \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            A->err = std::string(#call) + ": " + hipGetErrorString(e_);                           \
            return MCM_GPU_ERR_HIP;                                                               \
        }                                                                                         \
    } while (0)

template <typename T>
int adj_alloc(mcm_gpu_adjoint* A, std::vector<void*>& list, T** p, size_t n) {
    void* q = nullptr;
    if (n == 0) n = 1;
    HIPCHK_A(hipMalloc(&q, n * sizeof(T)));
    HIPCHK_A(hipMemsetAsync(q, 0, n * sizeof(T), A->ctx->stream));
    list.push_back(q);
    *p = (T*)q;
    return 0;
}

const size_t kSnapWords = SN_N;

double* slot_ptr(mcm_gpu_adjoint* A, bool ck, int s) {
    return (ck ? A->ckpt : A->seg) + (size_t)s * kSnapWords * A->np;
}

int adj_save(mcm_gpu_adjoint* A, bool ck, int s) {
    mcm_gpu_ctx* ctx = A->ctx;
    dim3 grid((unsigned)std::min<long>(blocks_for(A->np, MCM_BLOCK), 1024), SN_N);
    k_snap<<<grid, MCM_BLOCK, 0, ctx->stream>>>(ctx->d, slot_ptr(A, ck, s), (ck ? A->ckpt_sc : A->seg_sc) + s, 1);
    (ck ? A->ckpt_cur : A->seg_cur)[s] = ctx->cur;
    HIPCHK_A(hipGetLastError());
    return 0;
}

int adj_restore(mcm_gpu_adjoint* A, bool ck, int s) {
    mcm_gpu_ctx* ctx = A->ctx;
    dim3 grid((unsigned)std::min<long>(blocks_for(A->np, MCM_BLOCK), 1024), SN_N);
    k_snap<<<grid, MCM_BLOCK, 0, ctx->stream>>>(ctx->d, slot_ptr(A, ck, s), (ck ? A->ckpt_sc : A->seg_sc) + s, 0);
    ctx->cur = (ck ? A->ckpt_cur : A->seg_cur)[s];
    ctx->phase = 0;
    ctx->done = 0;
    ctx->fresh = false;
    HIPCHK_A(hipGetLastError());
    return 0;
}

// search slots for every step when they fit in an eighth of the free device
// memory and in 32 GB; otherwise the reverse sweep searches again
void adj_snap_alloc(mcm_gpu_adjoint* A) {
    const DevState& d = A->ctx->d;
    // slots rounded to 256 bytes: every part that is copied 16 bytes at a
    // time starts 16-byte aligned
    const size_t ni = ((size_t)d.npad * d.ncap + 4 * (size_t)d.np + 3 * (size_t)d.ngcap + 4 + 63) / 64 * 64;
    const size_t nd = ((size_t)d.np + (size_t)RS_SLOTS * 2 * ((size_t)d.np + d.ngcap) +
                       (size_t)RM_SLOTS * 2 * d.ngcap + SNAP_GD * (size_t)d.ngcap + 31) / 32 * 32;
    const size_t bytes = (size_t)A->nsteps * (ni * sizeof(int) + nd * sizeof(double) + sizeof(Scalars));
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return;
    if (bytes > freeb / 8 || bytes > (32ull << 30)) return;
    if (adj_alloc(A, A->lallocs, &A->snap_i, ni * A->nsteps) || adj_alloc(A, A->lallocs, &A->snap_d, nd * A->nsteps) ||
        adj_alloc(A, A->lallocs, &A->snap_sc, (size_t)A->nsteps)) {
        for (void* p : A->lallocs) (void)hipFree(p);
        A->lallocs.clear();
        A->snap_i = nullptr;
        A->snap_d = nullptr;
        A->snap_sc = nullptr;
        (void)hipGetLastError();
        A->err.clear();
        return;
    }
    A->snap_ni = ni;
    A->snap_nd = nd;
    A->snap_npad = d.npad;
    A->snap_ncap = d.ncap;
    A->snap_ngcap = d.ngcap;
}

bool adj_snap_fits(const mcm_gpu_adjoint* A) {
    const DevState& d = A->ctx->d;
    return A->snap_i && d.npad == A->snap_npad && d.ncap == A->snap_ncap && d.ngcap == A->snap_ngcap;
}

// the run's list array becomes step n's slot rows (same layout and size)
void adj_rows_in_slot(mcm_gpu_adjoint* A, int n) {
    if (!A->nbr_own) A->nbr_own = A->ctx->d.nbr;
    A->ctx->d.nbr = A->snap_i + (size_t)n * A->snap_ni;
}

// the run's own list array back (before anything that may free or grow it)
void adj_rows_home(mcm_gpu_adjoint* A) {
    if (A->nbr_own) {
        A->ctx->d.nbr = A->nbr_own;
        A->nbr_own = nullptr;
    }
}

struct RowsHome {
    mcm_gpu_adjoint* A;
    ~RowsHome() { adj_rows_home(A); }
};

// save (1) or restore (0) step n's search outputs
int adj_search(mcm_gpu_adjoint* A, int n, int save) {
    mcm_gpu_ctx* ctx = A->ctx;
    const DevState& d = ctx->d;
    SearchSlot p;
    int* pi = A->snap_i + (size_t)n * A->snap_ni;
    double* pd = A->snap_d + (size_t)n * A->snap_nd;
    const long np = d.np, gc = d.ngcap;
    p.rows = pi; pi += (size_t)d.npad * d.ncap;
    p.nreal = pi; pi += np;
    p.ntot = pi; pi += np;
    p.sorted = pi; pi += np;
    p.spos = pi; pi += np;
    p.gpmat = pi; pi += gc;
    p.gpar = pi; pi += gc;
    p.gsgn = pi; pi += gc;
    p.range = pi;
    p.rs = pd; pd += (size_t)RS_SLOTS * 2 * (np + gc);   // first: 16-byte aligned, as is rmg after it
    p.rmg = pd; pd += (size_t)RM_SLOTS * 2 * gc;
    p.mindist = pd; pd += np;
    p.gd = pd;
    const int nblk = std::max(1, std::min(blocks_for(RS_SLOTS * (np + gc), MCM_BLOCK), 512));
    k_adj_search<<<nblk, MCM_BLOCK, 0, ctx->stream>>>(d, p, A->snap_sc + n, save);
    HIPCHK_A(hipGetLastError());
    return 0;
}

// nsteps fused steps from the current state, the sampled positions after each
// step written to hist (device) when given; a step that outgrows a capacity
// makes the run grow and the call start again from checkpoint s, which holds
// its first state.  first: the index of the first step in the recording
// (whose list-build outputs are kept when the recording keeps them).
int adj_steps(mcm_gpu_adjoint* A, int nsteps, double* hist, int s, int first) {
    mcm_gpu_ctx* ctx = A->ctx;
    RowsHome home{A};
    int rc = 0;
    for (int attempt = 0; attempt < 16; attempt++) {
        Scalars s0;
        if ((rc = get_scalars_raw(ctx, &s0))) { A->err = ctx->err; return rc; }
        const bool keep = A->snap_ok && adj_snap_fits(A);
        if (!keep) A->snap_ok = false;
        for (int n = 0; n < nsteps; n++) {
            if (keep) {
                adj_rows_in_slot(A, first + n);
                const Target T = one(ctx);
                launch_neighbours(T, ctx->cur);
                if ((rc = adj_search(A, first + n, 1))) return rc;
                launch_strain(T, ctx->cur);
                launch_constitutive(T, ctx->cur);
                launch_momentum(T, ctx->cur);
            } else {
                launch_step(one(ctx), ctx->cur);
            }
            ctx->cur ^= 1;
            if (hist) k_adj_sample<<<blocks_for(A->nsample, 64), 64, 0, ctx->stream>>>(
                ctx->d, A->d_sample, A->nsample, hist + (size_t)n * 3 * A->nsample);
        }
        HIPCHK_A(hipGetLastError());
        Scalars s1;
        if ((rc = get_scalars_raw(ctx, &s1))) { A->err = ctx->err; return rc; }
        adj_rows_home(A);
        if (!s1.abort && s1.timestep - s0.timestep == nsteps) return MCM_GPU_OK;
        if ((rc = regrow(ctx, s1))) { A->err = ctx->err; return rc; }
        if ((rc = adj_restore(A, true, s))) return rc;
    }
    A->err = "steps did not fit after regrowing";
    return MCM_GPU_ERR_CAPACITY;
}

// ghost-sized adjoint buffers for the run's current capacities
int adj_ghost_buffers(mcm_gpu_adjoint* A) {
    const DevState& d = A->ctx->d;
    AdjDev& a = A->a;
    if (a.ngcap >= d.ngcap && a.gtcap >= d.ncap && !A->gallocs.empty()) return 0;
    for (void* p : A->gallocs) (void)hipFree(p);
    A->gallocs.clear();
    a.ngcap = d.ngcap;
    a.gtcap = std::max(a.gtcap, d.ncap);
    int rc = 0;
    for (int c = 0; c < 3; c++) {
        rc |= adj_alloc(A, A->gallocs, &a.gxb[c], a.ngcap);
        rc |= adj_alloc(A, A->gallocs, &a.gvb[c], a.ngcap);
    }
    rc |= adj_alloc(A, A->gallocs, &a.grb, a.ngcap);
    for (int c = 0; c < 6; c++) rc |= adj_alloc(A, A->gallocs, &a.gAb[c], a.ngcap);
    rc |= adj_alloc(A, A->gallocs, &a.gt_cnt, a.ngcap);
    rc |= adj_alloc(A, A->gallocs, &a.gt, (size_t)a.ngcap * a.gtcap);
    rc |= adj_alloc(A, A->gallocs, &a.gts, (size_t)a.ngcap * a.gtcap);
    return rc;
}

// intermediate slots for k steps at the run's current capacities
int adj_inter_buffers(mcm_gpu_adjoint* A) {
    const DevState& d = A->ctx->d;
    const long ucap = (long)d.np + d.ngcap, nbrlen = (long)d.npad * d.ncap;
    if (!A->iallocs.empty() && A->lay.ucap == ucap && A->lay.nbrlen == nbrlen) return 0;
    for (void* p : A->iallocs) (void)hipFree(p);
    A->iallocs.clear();
    A->lay = InterLayout{ucap, (long)d.np, nbrlen, (long)d.ngcap};
    A->inter_dn = (size_t)(INTER_UCAP * ucap + 13 * (long)d.np);
    A->inter_in = (size_t)(nbrlen + 3 * (long)d.np + 2 * (long)d.ngcap);
    int rc = 0;
    rc |= adj_alloc(A, A->iallocs, &A->inter_d, A->inter_dn * A->ninter);
    rc |= adj_alloc(A, A->iallocs, &A->inter_i, A->inter_in * A->ninter);
    rc |= adj_alloc(A, A->iallocs, &A->inter_sc, A->ninter);
    return rc;
}

int adj_inter(mcm_gpu_adjoint* A, int s, bool save) {
    mcm_gpu_ctx* ctx = A->ctx;
    const long big = std::max(A->lay.nbrlen, INTER_UCAP * A->lay.ucap);
    dim3 grid((unsigned)std::min<long>(blocks_for(big / 8 + 1, MCM_BLOCK), 1024), INTER_ND + INTER_NI);
    const int skip_rows = A->snap_ok && adj_snap_fits(A) && A->nbr_own != nullptr;
    k_inter<<<grid, MCM_BLOCK, 0, ctx->stream>>>(ctx->d, A->lay, A->inter_d + (size_t)s * A->inter_dn,
                                                 A->inter_i + (size_t)s * A->inter_in, A->inter_sc + s, save ? 1 : 0,
                                                 skip_rows);
    HIPCHK_A(hipGetLastError());
    return 0;
}

// the stages after the neighbour search, in the reference's call sequence
int adj_granular_rest(mcm_gpu_adjoint* A) {
    mcm_gpu_ctx* ctx = A->ctx;
    int rc = 0;
    if (!rc) rc = mcm_gpu_strain(ctx);
    if (!rc) rc = mcm_gpu_update_h(ctx);
    if (!rc) rc = mcm_gpu_constitutive(ctx);
    if (!rc) rc = mcm_gpu_dt_update(ctx);
    if (!rc) rc = mcm_gpu_momentum(ctx);
    if (!rc) rc = mcm_gpu_boundcond(ctx);
    if (!rc) rc = mcm_gpu_update_velocity(ctx);
    if (!rc) rc = mcm_gpu_move(ctx);
    if (rc) A->err = ctx->err;
    return rc;
}

// one recomputed step of a segment: the fused step's kernels with their
// diagnostics kept (rod, spin, critts, the acceleration after the boundary
// conditions, rhoold, etry, einc), which leave on the device every
// intermediate the reverse step reads, as the granular stages do, in four
// pair and particle launches instead of eleven (plain SPH, fixed h, no
// contact or XSPH: the adjoint's scope)
int adj_fused_diag_step(mcm_gpu_adjoint* A, int n) {
    mcm_gpu_ctx* ctx = A->ctx;
    const DevState& d = ctx->d;
    hipStream_t st = ctx->stream;
    const int cur = ctx->cur;
    if (A->snap_ok && adj_snap_fits(A)) {   // the recording's search of step n
        adj_rows_in_slot(A, n);
        int rc = adj_search(A, n, 0);
        if (rc) return rc;
    } else {
        launch_neighbours(one(ctx), cur);
        k_adj_check<<<1, 1, 0, st>>>(ctx->d, A->a);
    }
    k_strain_const<MODE_FSTRAIN, 0, false><<<ctx->nb, MCM_BLOCK, 0, st>>>(ctx->d, cur);
    k_strain_const<MODE_FCONST, 0, true><<<ctx->nb, MCM_BLOCK, 0, st>>>(ctx->d, cur);
    k_strain_fold<MODE_FCONST><<<1, MCM_BLOCK, 0, st>>>(ctx->d);
    if (d.boundary) k_ghost_stress<<<ghost_blocks(d), MCM_BLOCK, 0, st>>>(ctx->d, cur);
    k_momentum<MODE_FUSED, 0, true><<<ctx->nb, MCM_BLOCK, 0, st>>>(ctx->d, cur);
    launch_c_fold(one(ctx), 1);
    HIPCHK_A(hipGetLastError());
    ctx->cur ^= 1;
    return 0;
}

// recording with every step kept: each step runs through the granular stages
// (the neighbour search with its capacity check) and its start state and
// intermediates are saved; a capacity that grows changes the slot layout, so
// the recording starts again from the initial state
int adj_record_store(mcm_gpu_adjoint* A) {
    mcm_gpu_ctx* ctx = A->ctx;
    const DevState& d = ctx->d;
    int rc = 0;
    for (int attempt = 0; attempt < 8; attempt++) {
        if (attempt > 0 && (rc = adj_restore(A, true, 0))) return rc;
        if ((rc = adj_inter_buffers(A))) return rc;
        bool grown = false;
        for (int n = 0; n < A->nsteps && !grown; n++) {
            if ((rc = adj_save(A, true, n))) return rc;
            const int ncap = d.ncap, ngcap = d.ngcap, cellcap = d.cellcap;
            if ((rc = mcm_gpu_neighbours(ctx))) { A->err = ctx->err; return rc; }
            if (d.ncap != ncap || d.ngcap != ngcap || d.cellcap != cellcap) { grown = true; break; }
            if ((rc = adj_granular_rest(A))) return rc;
            if ((rc = adj_inter(A, n, true))) return rc;
            if (A->nsample > 0)
                k_adj_sample<<<blocks_for(A->nsample, 64), 64, 0, ctx->stream>>>(
                    ctx->d, A->d_sample, A->nsample, A->d_hist + (size_t)n * 3 * A->nsample);
        }
        if (!grown) {
            HIPCHK_A(hipGetLastError());
            return 0;
        }
    }
    A->err = "recording did not fit after regrowing";
    return MCM_GPU_ERR_CAPACITY;
}

// reverse of the step just recomputed (start-of-step state in slot (ck, s))
int adj_reverse_step(mcm_gpu_adjoint* A, bool ck, int s) {
    mcm_gpu_ctx* ctx = A->ctx;
    const DevState& d = ctx->d;
    AdjDev& a = A->a;
    hipStream_t st = ctx->stream;
    int rc = adj_ghost_buffers(A);
    if (rc) return rc;
    a.snap = slot_ptr(A, ck, s);
    a.cur = (ck ? A->ckpt_cur : A->seg_cur)[s];
    const int nb = ctx->nb, gb = ghost_blocks(d);
    const int wb = blocks_for(A->np, MCM_BLOCK / kLpp);                                  // kLpp lanes per particle
    const int wg = std::max(1, std::min(blocks_for(d.ngcap, MCM_BLOCK / 64), 4096));   // wave per ghost
    // transposed lists; an overflow sets err bit 1 (checked after the sweep)
    AdjTiming* T = A->timing;
    int tk = 0;
    auto mark = [&]() {
        if (T) (void)hipEventRecord(T->ev[tk++], st);
    };
    mark();
    k_adj_zero_counts<<<std::max(1, std::min(blocks_for(std::max((long)a.ngcap, A->np), MCM_BLOCK), 1024)), MCM_BLOCK, 0,
                        st>>>(a, a.ngcap, A->np);
    mark();
    const int nbf = blocks_for(A->np, (MCM_BLOCK / 64) * ADJ_GT_PPW);
    k_adj_fill<<<nbf + gb, MCM_BLOCK, 0, st>>>(d, a, nbf);
    mark();
    k_adj_sort<<<wg + nb, MCM_BLOCK, 0, st>>>(d, a, wg);
    mark();
    k_adj_move<<<nb, MCM_BLOCK, 0, st>>>(d, a);
    mark();
    k_adj_fold_move<<<1, MCM_BLOCK, 0, st>>>(d, a, nb);
    mark();
    k_adj_mom<<<wb, MCM_BLOCK, 0, st>>>(d, a);
    mark();
    k_adj_mom_ghost<<<wg, MCM_BLOCK, 0, st>>>(d, a);
    mark();
    k_adj_gmap_mom<<<nb, MCM_BLOCK, 0, st>>>(d, a);
    mark();
    k_adj_const<<<dim3(nb, ADJ_NPASS), MCM_BLOCK, 0, st>>>(d, a);
    mark();
    k_adj_const_fin<<<nb, MCM_BLOCK, 0, st>>>(d, a);
    mark();
    k_adj_strain<<<wb, MCM_BLOCK, 0, st>>>(d, a);
    mark();
    k_adj_strain_ghost<<<wg, MCM_BLOCK, 0, st>>>(d, a);
    mark();
    k_adj_gmap_strain<<<nb, MCM_BLOCK, 0, st>>>(d, a);
    mark();
    k_adj_fold_end<<<1, MCM_BLOCK, 0, st>>>(d, a, nb);
    mark();
    HIPCHK_A(hipGetLastError());
    if (T) {
        HIPCHK_A(hipEventSynchronize(T->ev[ADJ_TK]));
        for (int k = 0; k < ADJ_TK; k++) {
            float ms = 0.0f;
            HIPCHK_A(hipEventElapsedTime(&ms, T->ev[k], T->ev[k + 1]));
            T->ms[k] += ms;
        }
        T->steps++;
    }
    return 0;
}

// Parameter adjoints from the initial state (host; mcm_initial.f90:33-101):
// sigma_0 = -p_0 I, rho_0 and e_0 per particle, and dt_0 = tssf min critts_0
// (critts of the argmin particle, which with tcrit_opt 1 also reads |v_0|).
// The initial state's derivatives come from init_particle with dual material
// constants; the last dual direction is |v_0|.
void adj_initial_params(mcm_gpu_adjoint* A, const std::vector<double> sb[6], const std::vector<double>& rb,
                        const std::vector<double>& eb, double dtb, double* pbar, double* v0bar,
                        const std::vector<double>& v0) {
    typedef Dual<ADJ_NPAR + 1> D;
    mcm_gpu_ctx* ctx = A->ctx;
    const HostDeck& Dk = ctx->deck;
    double cmin = 1.0e300;
    long imin = -1;
    D cbest;
    for (long i = 0; i < A->np; i++) {
        const int m = Dk.pmat[i];
        const MatParams& M = Dk.mat[m];
        MatParamsT<D> Md;
        Md.model = M.model; Md.eos = M.eos; Md.visc_type = M.visc_type; Md.ispall = M.ispall;
        Md.efps_flag = M.efps_flag; Md.pad_ = 0;
        Md.rho = M.rho; Md.g = M.g; Md.av_l = M.av_l; Md.av_q = M.av_q; Md.h = M.h; Md.rho_max = M.rho_max;
        for (int k = 0; k < 48; k++) Md.strinput[k] = M.strinput[k];
        for (int k = 0; k < 24; k++) Md.eosinput[k] = M.eosinput[k];
        if (M.model == 10) {
            Md.strinput[0].d[0] = 1.0;
            Md.strinput[1].d[1] = 1.0;
            Md.strinput[2].d[2] = 1.0;
            Md.eosinput[1].d[3] = 1.0;
            Md.g = Md.strinput[0];
        } else {   // model 1: E, nu and the constants derived from them
            const MatParams& R = Dk.mat_raw[m];
            D E = R.strinput[0], nu = R.strinput[5];
            E.d[0] = 1.0;
            nu.d[1] = 1.0;
            model1_constants(Md, E, nu);
        }
        PStateT<D> s{};
        init_particle(s, Md, Dk, i, Dk.pm[i], M.strinput[3]);
        s.tracerod = 0.0;
        s.h = Dk.init_h_opt == 1 ? Dk.ph[i] : M.h;   // mcm_init_h.f90:22-48
        double va = 0.0;
        for (int c = 0; c < 3; c++) va = va + v0[3 * i + c] * v0[3 * i + c];
        s.vabs = std::sqrt(va);
        s.vabs.d[ADJ_NPAR] = 1.0;
        const D cr = critts(s, Md, Dk.tcrit_opt, 0.0);
        const double sp = -(sb[0][i] + sb[1][i] + sb[2][i]);
        for (int k = 0; k < ADJ_NPAR; k++)
            pbar[ADJ_NPAR * m + k] += sp * s.p.d[k] + rb[i] * s.rho.d[k] + eb[i] * s.e.d[k];
        if (cr.v < cmin) { cmin = cr.v; imin = i; cbest = cr; }
    }
    const bool itss_binds = Dk.itss != 0.0 && Dk.itss < cmin * Dk.tssf;
    if (imin >= 0 && !itss_binds) {
        const int m = Dk.pmat[imin];
        for (int k = 0; k < ADJ_NPAR; k++) pbar[ADJ_NPAR * m + k] += dtb * Dk.tssf * cbest.d[k];
        const double va = std::sqrt(v0[3 * imin] * v0[3 * imin] + v0[3 * imin + 1] * v0[3 * imin + 1] +
                                                         v0[3 * imin + 2] * v0[3 * imin + 2]);
        if (va > 0.0)
            for (int c = 0; c < 3; c++)
                v0bar[3 * imin + c] += dtb * Dk.tssf * cbest.d[ADJ_NPAR] * v0[3 * imin + c] / va;
    }
}

static int adj_sweep(mcm_gpu_adjoint* A, const double* xbar, const double* vbar, const double* histbar,
                     double* v0bar, double* pbar, int* flags);

// the recording itself (buffers allocated): steps from the run's current
// state, checkpoints, sampled positions to hist
int adj_record(mcm_gpu_adjoint* A, double* hist) {
    mcm_gpu_ctx* ctx = A->ctx;
    const int nsteps = A->nsteps, nsample = A->nsample;
    int rc = 0;
    if (A->store) {
        if ((rc = adj_record_store(A))) return rc;
    } else {
        A->snap_ok = adj_snap_fits(A);
        for (int sgi = 0; sgi < A->nseg; sgi++) {
            const int s0 = sgi * A->k, s1 = std::min(s0 + A->k, nsteps);
            if ((rc = adj_save(A, true, sgi))) return rc;
            if ((rc = adj_steps(A, s1 - s0, nsample > 0 ? A->d_hist + (size_t)s0 * 3 * nsample : nullptr, sgi, s0)))
                return rc;
        }
        // the last step's lists (built in its slot) into the run's own array,
        // which describes the run's state after the recording
        if (A->snap_ok && adj_snap_fits(A))
            HIPCHK_A(hipMemcpyAsync(ctx->d.nbr, A->snap_i + (size_t)(nsteps - 1) * A->snap_ni,
                                    sizeof(int) * (size_t)ctx->d.npad * ctx->d.ncap, hipMemcpyDeviceToDevice,
                                    ctx->stream));
    }
    if (nsample > 0)
        HIPCHK_A(hipMemcpyAsync(hist, A->d_hist, sizeof(double) * 3 * nsample * (size_t)nsteps,
                                hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK_A(hipStreamSynchronize(ctx->stream));
    return MCM_GPU_OK;
}

}  // namespace

extern "C" {

int mcm_gpu_set_material_params(mcm_gpu_ctx* ctx, int32_t mat, const double* params) {
    if (!ctx || !params || mat < 0 || mat >= ctx->deck.nummat) return MCM_GPU_ERR_ARG;
    MatParams R = ctx->deck.mat_raw[mat];
    if (R.model == 10) {
        R.strinput[0] = params[0];
        R.strinput[1] = params[1];
        R.strinput[2] = params[2];
        R.eosinput[1] = params[3];
    } else if (R.model == 1) {
        R.strinput[0] = params[0];   // E (card 1)
        R.strinput[5] = params[1];   // nu (card 2)
    } else {
        ctx->err = "material parameters are defined for models 1 and 10";
        return MCM_GPU_ERR_MODEL;
    }
    ctx->deck.mat_raw[mat] = R;
    setup_material(R);   // derived constants as the reference's set-up routines leave them
    ctx->deck.mat[mat] = R;
    return MCM_GPU_OK;
}

int mcm_gpu_get_material_params(mcm_gpu_ctx* ctx, int32_t mat, double* params) {
    if (!ctx || !params || mat < 0 || mat >= ctx->deck.nummat) return MCM_GPU_ERR_ARG;
    const MatParams& R = ctx->deck.mat_raw[mat];
    if (R.model == 1) {
        params[0] = R.strinput[0];
        params[1] = R.strinput[5];
        params[2] = 0.0;
        params[3] = 0.0;
    } else {
        params[0] = R.strinput[0];
        params[1] = R.strinput[1];
        params[2] = R.strinput[2];
        params[3] = R.eosinput[1];
    }
    return MCM_GPU_OK;
}

int mcm_gpu_set_initial_velocity(mcm_gpu_ctx* ctx, const double* v, int64_t n) {
    if (!ctx || !v || n != 3 * (int64_t)ctx->deck.np) return MCM_GPU_ERR_ARG;
    ctx->deck.v.assign(v, v + n);
    return MCM_GPU_OK;
}

int mcm_gpu_adj_forward(mcm_gpu_ctx* ctx, int32_t nsteps, int32_t ckpt_every, const int32_t* sample,
                        int32_t nsample, double* hist, mcm_gpu_adjoint** out) {
    if (!ctx || !out || nsteps <= 0 || ckpt_every <= 0 || nsample < 0 || (nsample > 0 && (!sample || !hist)))
        return MCM_GPU_ERR_ARG;
    mcm_gpu_adjoint* prev = *out;   // a recording of this run whose buffers may be reused, or NULL
    if (prev && prev->ctx != ctx) { ctx->err = "the recording to reuse belongs to another run"; return MCM_GPU_ERR_ARG; }
    const DevState& d = ctx->d;
    std::string why;
    if (ctx->phase != 0 || ctx->done != 0 || ctx->batch)
        why = "the run must be at a step boundary and not part of a batch";
    else if (d.disctype != 0) why = "the adjoint covers plain SPH (disctype 0)";
    else if (d.h_opt != 0) why = "the adjoint covers fixed h (h_opt 0)";
    else if (d.tcrit_opt != 0 && d.tcrit_opt != 1) why = "the adjoint covers tcrit_opt 0 and 1";
    else if (d.contacttype > 0 || d.veloc_opt == 1 || d.drelax || d.baseaccel)
        why = "the adjoint does not cover penalty contact, XSPH, dynamic relaxation or base acceleration";
    else if (d.skin > 0.0) why = "the adjoint covers runs without a neighbour skin";
    for (int m = 0; m < d.nummat && why.empty(); m++)
        if (ctx->deck.mat[m].model != 10 && ctx->deck.mat[m].model != 1)
            why = "the adjoint covers material models 1 and 10";
    std::vector<char> seen(d.np, 0);
    for (int k = 0; k < nsample && why.empty(); k++) {
        if (sample[k] < 0 || sample[k] >= d.np || seen[sample[k]]) why = "sampled particles must be distinct ids below np";
        else seen[sample[k]] = 1;
    }
    if (!why.empty()) { ctx->err = why; return MCM_GPU_ERR_ARG; }
    ctx->fresh = false;   // the recorded steps are fused steps, each with its own search
    int rc = 0;
    if (prev && prev->nsteps == nsteps && prev->k == ckpt_every && prev->nsample == nsample && prev->np == d.np) {
        // same shape: the device buffers are reused (no allocation, no free,
        // which would synchronise the device)
        mcm_gpu_adjoint* A = prev;
        A->err.clear();
        if (nsample > 0)
            HIPCHK_A(hipMemcpyAsync(A->d_sample, sample, sizeof(int) * nsample, hipMemcpyHostToDevice, ctx->stream));
        return adj_record(A, hist);
    }
    if (prev) mcm_gpu_adj_destroy(prev);
    mcm_gpu_adjoint* A = new mcm_gpu_adjoint();
    *out = A;
    A->ctx = ctx;
    ctx->recs.push_back(A);
    A->nsteps = nsteps;
    A->k = ckpt_every;
    if (const char* e = getenv("MCM_ADJ_TIMING")) {
        if (e[0] == '1') {
            A->timing = new AdjTiming();
            for (hipEvent_t& ev : A->timing->ev)
                if (hipEventCreate(&ev) != hipSuccess) {
                    A->err = "MCM_ADJ_TIMING: hipEventCreate failed";
                    return MCM_GPU_ERR_HIP;
                }
        }
    }
    A->nseg = (nsteps + ckpt_every - 1) / ckpt_every;
    A->store = ckpt_every == 1;
    A->ninter = A->store ? nsteps : ckpt_every;
    A->nsample = nsample;
    A->np = d.np;
    A->ckpt_cur.assign(A->nseg, 0);
    A->seg_cur.assign(ckpt_every, 0);
    const size_t slot = kSnapWords * A->np;
    rc |= adj_alloc(A, A->allocs, &A->ckpt, slot * A->nseg);
    rc |= adj_alloc(A, A->allocs, &A->ckpt_sc, A->nseg);
    rc |= adj_alloc(A, A->allocs, &A->d_sample, std::max(nsample, 1));
    rc |= adj_alloc(A, A->allocs, &A->d_hist, (size_t)nsteps * 3 * std::max(nsample, 1));
    AdjDev& a = A->a;
    const size_t n = A->np;
    for (int c = 0; c < 3; c++) {
        rc |= adj_alloc(A, A->allocs, &a.xb[c], n);
        rc |= adj_alloc(A, A->allocs, &a.vb[c], n);
    }
    rc |= adj_alloc(A, A->allocs, &a.abp, 4 * n);
    for (int c = 0; c < 6; c++) rc |= adj_alloc(A, A->allocs, &a.sb[c], n);
    rc |= adj_alloc(A, A->allocs, &a.lbp, 10 * n);
    rc |= adj_alloc(A, A->allocs, &a.xin, n * ADJ_NIN);
    rc |= adj_alloc(A, A->allocs, &a.rb, n);
    rc |= adj_alloc(A, A->allocs, &a.eb, n);
    rc |= adj_alloc(A, A->allocs, &a.fb, n);
    rc |= adj_alloc(A, A->allocs, &a.qb, n);
    rc |= adj_alloc(A, A->allocs, &a.rbn, n);
    rc |= adj_alloc(A, A->allocs, &a.pdt, n);
    rc |= adj_alloc(A, A->allocs, &a.gof_cnt, n);
    rc |= adj_alloc(A, A->allocs, &a.gof, n * ADJ_GOF);
    rc |= adj_alloc(A, A->allocs, &a.red, (size_t)ctx->nb * RA_N + 64);
    rc |= adj_alloc(A, A->allocs, &a.as, AS_N);
    rc |= adj_alloc(A, A->allocs, &a.err, 1);
    if (rc) return rc;
    if (!A->store) adj_snap_alloc(A);
    a.gtcap = 0;
    a.ngcap = 0;
    if (nsample > 0)
        HIPCHK_A(hipMemcpyAsync(A->d_sample, sample, sizeof(int) * nsample, hipMemcpyHostToDevice, ctx->stream));
    return adj_record(A, hist);
}

int mcm_gpu_adj_backward(mcm_gpu_adjoint* A, const double* xbar, const double* vbar, const double* histbar,
                         double* v0bar, double* pbar) {
    if (!A || !v0bar || !pbar) return MCM_GPU_ERR_ARG;
    for (int attempt = 0; attempt < 4; attempt++) {
        int flags = 0;
        const int rc = adj_sweep(A, xbar, vbar, histbar, v0bar, pbar, &flags);
        if (rc) return rc;
        if (flags & 2) { A->err = "a recomputed step outgrew a capacity"; return MCM_GPU_ERR_CAPACITY; }
        if (!(flags & 1)) return MCM_GPU_OK;
        // more real partners of one ghost than transposed slots: sweep again
        A->a.gtcap *= 2;
        for (void* p : A->gallocs) (void)hipFree(p);
        A->gallocs.clear();
    }
    A->err = "transposed ghost lists overflowed";
    return MCM_GPU_ERR_CAPACITY;
}

}  // extern "C"

namespace {

static int adj_sweep(mcm_gpu_adjoint* A, const double* xbar, const double* vbar, const double* histbar, double* v0bar,
              double* pbar, int* flags) {
    mcm_gpu_ctx* ctx = A->ctx;
    hipStream_t st = ctx->stream;
    AdjDev& a = A->a;
    const long np = A->np;
    RowsHome home{A};
    int rc = 0;
    // adjoint of the end state (the run's stream is non-blocking: plain
    // copies are ordered with it by synchronising it)
    HIPCHK_A(hipStreamSynchronize(st));
    std::vector<double> h(np);
    for (int c = 0; c < 3; c++) {
        for (long i = 0; i < np; i++) h[i] = xbar ? xbar[3 * i + c] : 0.0;
        HIPCHK_A(hipMemcpy(a.xb[c], h.data(), sizeof(double) * np, hipMemcpyHostToDevice));
        for (long i = 0; i < np; i++) h[i] = vbar ? vbar[3 * i + c] : 0.0;
        HIPCHK_A(hipMemcpy(a.vb[c], h.data(), sizeof(double) * np, hipMemcpyHostToDevice));
    }
    double* zero[] = {a.rb, a.eb, a.fb, a.qb, a.rbn, a.sb[0], a.sb[1], a.sb[2], a.sb[3], a.sb[4], a.sb[5]};
    for (double* z : zero) HIPCHK_A(hipMemsetAsync(z, 0, sizeof(double) * np, st));
    HIPCHK_A(hipMemsetAsync(a.as, 0, sizeof(double) * AS_N, st));
    HIPCHK_A(hipMemsetAsync(a.err, 0, sizeof(int), st));
    const bool seeds = histbar && A->nsample > 0;
    if (seeds)
        HIPCHK_A(hipMemcpyAsync(A->d_hist, histbar, sizeof(double) * 3 * A->nsample * (size_t)A->nsteps,
                                hipMemcpyHostToDevice, st));
    if (A->store && (A->lay.ucap != (long)ctx->d.np + ctx->d.ngcap || A->lay.nbrlen != (long)ctx->d.npad * ctx->d.ncap)) {
        A->err = "the run's capacities changed after the recording";
        return MCM_GPU_ERR_ARG;
    }
    if (!A->store && !A->seg) {   // the segment's start states (first reverse sweep)
        rc |= adj_alloc(A, A->allocs, &A->seg, kSnapWords * A->np * A->k);
        rc |= adj_alloc(A, A->allocs, &A->seg_sc, A->k);
        if (rc) return rc;
    }
    if ((rc = adj_inter_buffers(A))) return rc;
    if ((rc = adj_ghost_buffers(A))) return rc;
    if (A->store) {   // every step's intermediates were kept by the recording
        for (int n = A->nsteps - 1; n >= 0; n--) {
            if ((rc = adj_inter(A, n, false))) return rc;
            if (seeds)
                k_adj_seed<<<blocks_for(A->nsample, 64), 64, 0, st>>>(a, A->d_sample, A->nsample,
                                                                      A->d_hist + (size_t)n * 3 * A->nsample);
            if ((rc = adj_reverse_step(A, true, n))) return rc;
        }
    }
    for (int sgi = A->store ? -1 : A->nseg - 1; sgi >= 0; sgi--) {
        const int s0 = sgi * A->k, s1 = std::min(s0 + A->k, A->nsteps);
        // recompute the segment with the fused kernels, keeping the start of
        // every step and the step's intermediates
        if ((rc = adj_restore(A, true, sgi))) return rc;
        for (int n = s0; n < s1; n++) {
            if ((rc = adj_save(A, false, n - s0))) return rc;
            if ((rc = adj_fused_diag_step(A, n))) return rc;
            if ((rc = adj_inter(A, n - s0, true))) return rc;
        }
        for (int n = s1 - 1; n >= s0; n--) {
            if (A->snap_ok && adj_snap_fits(A)) adj_rows_in_slot(A, n);   // step n's lists, in its slot
            if ((rc = adj_inter(A, n - s0, false))) return rc;
            if (seeds)
                k_adj_seed<<<blocks_for(A->nsample, 64), 64, 0, st>>>(a, A->d_sample, A->nsample,
                                                                      A->d_hist + (size_t)n * 3 * A->nsample);
            if ((rc = adj_reverse_step(A, false, n - s0))) return rc;
        }
    }
    // results
    HIPCHK_A(hipMemcpyAsync(flags, a.err, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK_A(hipStreamSynchronize(st));
    if (*flags) return MCM_GPU_OK;
    std::vector<double> vb[3], sbh[6], rbh(np), ebh(np);
    for (int c = 0; c < 3; c++) {
        vb[c].resize(np);
        HIPCHK_A(hipMemcpy(vb[c].data(), a.vb[c], sizeof(double) * np, hipMemcpyDeviceToHost));
    }
    for (int c = 0; c < 6; c++) {
        sbh[c].resize(np);
        HIPCHK_A(hipMemcpy(sbh[c].data(), a.sb[c], sizeof(double) * np, hipMemcpyDeviceToHost));
    }
    HIPCHK_A(hipMemcpy(rbh.data(), a.rb, sizeof(double) * np, hipMemcpyDeviceToHost));
    HIPCHK_A(hipMemcpy(ebh.data(), a.eb, sizeof(double) * np, hipMemcpyDeviceToHost));
    std::vector<double> as(AS_N);
    HIPCHK_A(hipMemcpy(as.data(), a.as, sizeof(double) * AS_N, hipMemcpyDeviceToHost));
    const HostDeck& Dk = ctx->deck;
    static const int mask[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
    // v_0 as the initialisation leaves it (mcm_init_boundcond.f90:25-54 zeroes
    // fixed components)
    std::vector<double> v0(3 * np);
    for (long i = 0; i < np; i++) {
        const int bc = Dk.dispbc[i];
        for (int c = 0; c < 3; c++) {
            const bool fixed = bc >= 1 && bc <= 7 && mask[bc][c];
            v0[3 * i + c] = fixed ? 0.0 : Dk.v[3 * i + c];
            v0bar[3 * i + c] = vb[c][i];
        }
    }
    const int nm = ctx->d.nummat;
    for (int k = 0; k < ADJ_NPAR * nm; k++) pbar[k] = as[AS_PAR + k];
    adj_initial_params(A, sbh, rbh, ebh, as[AS_DTB], pbar, v0bar, v0);
    for (long i = 0; i < np; i++) {
        const int bc = Dk.dispbc[i];
        for (int c = 0; c < 3; c++)
            if (bc >= 1 && bc <= 7 && mask[bc][c]) v0bar[3 * i + c] = 0.0;
    }
    return MCM_GPU_OK;
}

}  // namespace

extern "C" {

const char* mcm_gpu_adj_last_error(const mcm_gpu_adjoint* A) { return A ? A->err.c_str() : "null adjoint"; }

void mcm_gpu_adj_destroy(mcm_gpu_adjoint* A) {
    if (!A) return;
    if (A->ctx) {
        std::vector<mcm_gpu_adjoint*>& r = A->ctx->recs;
        r.erase(std::remove(r.begin(), r.end(), A), r.end());
        if (A->ctx->stream) (void)hipStreamSynchronize(A->ctx->stream);
    }
    for (void* p : A->allocs) (void)hipFree(p);
    for (void* p : A->gallocs) (void)hipFree(p);
    for (void* p : A->iallocs) (void)hipFree(p);
    for (void* p : A->lallocs) (void)hipFree(p);
    if (AdjTiming* T = A->timing) {
        if (T->steps > 0) {
            double tot = 0.0;
            for (int k = 0; k < ADJ_TK; k++) tot += T->ms[k];
            fprintf(stderr, "adjoint timing: %d steps, checkpoint every %d, %ld reverse steps, %.1f us of reverse kernels per step\n",
                    A->nsteps, A->k, T->steps, 1e3 * tot / T->steps);
            for (int k = 0; k < ADJ_TK; k++)
                fprintf(stderr, "adjoint timing: %-20s %9.2f us per step %6.1f %%\n", adj_tk_names[k],
                        1e3 * T->ms[k] / T->steps, 100.0 * T->ms[k] / tot);
        }
        for (hipEvent_t ev : T->ev)
            if (ev) (void)hipEventDestroy(ev);
        delete T;
    }
    delete A;
}

}  // extern "C"
