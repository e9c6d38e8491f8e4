// mcm_wave.hip -- wavefront-per-particle kernels of the fused MCM step.
//
// A 64-wide wavefront owns 64 consecutive particles and walks them one at a
// time.  For each particle the lanes take its neighbours (lane l holds list
// slots l and l + 64), load each neighbour's gather record once and keep the
// pair quantities (kernel value, gradient, volume, position and velocity or
// stress) in registers across the reference's three corrected passes, so a
// record is fetched once per kernel instead of once per pass.
//
// Every per-particle sum of the reference is a sequential fold over its list
// in list order.  The lanes write their terms to LDS and lane c adds up
// component c in list order, so each sum is rounded exactly as the
// reference's loop rounds it: results equal the thread-per-particle kernels
// (and, for plain SPH, the reference) bit for bit.
//
// The per-particle epilogues (update_h, constitutive model, EOS, critts;
// boundcond, energies, velocity, move) then run one particle per lane, as in
// the thread-per-particle kernels, and share their code.
//
// Lists of more than 2 * 64 entries are left to the thread-per-particle
// kernels (the host picks the kernel from the list capacity).
#pragma once
#include "mcm_step.hip"

#define WAVE_SLOTS 2
#define OSUM_ROW 66                          // doubles per component row (16-byte aligned rows)
#define OSUM_MAXC 9                          // components per ordered-sum round
#define OSUM_WORDS (OSUM_MAXC * OSUM_ROW + 16)

// LDS writes of one lane become visible to the other lanes of its wavefront
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// value of v in lane src (src uniform)
__device__ __forceinline__ double lane_value(double v, int src) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Ordered sums of S components over list entries 0..n-1.  t[s][c] is this
// lane's term of component c for list slot lane + 64 s.  Lane c folds
// component c in list order starting from 0.0, as the reference's loop does;
// the S sums are returned in every lane.
template <int S>
__device__ __forceinline__ void ordered_sums(double* buf, const double (&t)[WAVE_SLOTS][S], int n, double (&out)[S]) {
    static_assert(S <= OSUM_MAXC, "ordered_sums: too many components");
    const int lane = threadIdx.x & 63;
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < WAVE_SLOTS; s++) {
        const int kn = min(n - 64 * s, 64);
        if (kn <= 0) break;
#pragma unroll
        for (int c = 0; c < S; c++) buf[c * OSUM_ROW + lane] = t[s][c];
        wave_lds_sync();
        if (lane < S) {
            const double* row = buf + lane * OSUM_ROW;
            int k = 0;
            for (; k + 2 <= kn; k += 2) {
                double2 v = *reinterpret_cast<const double2*>(row + k);
                acc = acc + v.x;
                acc = acc + v.y;
            }
            if (k < kn) acc = acc + row[k];
        }
        wave_lds_sync();
    }
    if (lane < S) buf[OSUM_MAXC * OSUM_ROW + lane] = acc;
    wave_lds_sync();
#pragma unroll
    for (int c = 0; c < S; c++) out[c] = buf[OSUM_MAXC * OSUM_ROW + c];
    wave_lds_sync();
}

// lane p keeps component values of particle p
template <int S>
__device__ __forceinline__ void keep_if_lane(int p, const double (&v)[S], double (&mine)[S]) {
    const bool me = (threadIdx.x & 63) == p;
#pragma unroll
    for (int c = 0; c < S; c++) mine[c] = me ? v[c] : mine[c];
}

// ------------------------------------------------------------ list build
// mcm_ll_neighbours.f90:96-299, one wavefront per particle.  Lanes 0..26
// look up the 27 cells around the particle's cell (z, y, x order); the
// candidates of each cell run (a contiguous x-run of cells, or single cells
// where a row holds ghost-headed cells) are tested 64 at a time, and a
// ballot with a prefix count packs the accepted ids into the list row in
// walk order.
__global__ __launch_bounds__(256) void k_list_wave(DevState d_, int cur) {
    Blk bk;
    const DevState& d = enter<true>(d_, bk);
    const int lb = bk.b;
    Scalars* sc = d.sc;
    if (sc->abort) return;
    const int lane = threadIdx.x & 63;
    const int wbase = lb * 256 + wave_uniform(threadIdx.x >> 6) * 64;
    const GridDev g = sc->grid;
    const int L0 = g.lim[0], L1 = g.lim[1], L2 = g.lim[2];
    const int ncap = d.ncap, np = d.np;
    const int npass = d.nummat > 1 ? 2 : 1;
    const double2* cd = reinterpret_cast<const double2*>(d.cand);
    const int2* ci = reinterpret_cast<const int2*>(d.cand_id);
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int wmax = 0;
    for (int p = 0; p < 64; p++) {
        const int i = wbase + p;
        if (i >= np) break;
        int* out = nbr_row(d, d.spos[i]);
        int cnt = 0;
        double mind = 1.0e20, mind_cp = 1.0e20;
        const int c = d.cell_of[i];
        int nreal;
        if (c >= 0) {
            const double xi = d.x0[0][i], yi = d.x0[1][i], zi = d.x0[2][i];
            const double hi = d.hbuf[cur][i];
            const int mi = d.pmat[i];
            const int b0 = c % L0, b1 = (c / L0) % L1, b2 = c / (L0 * L1);
            // lane t < 27 <-> cell (b2 - 1 + t / 9, b1 - 1 + (t / 3) % 3, b0 - 1 + t % 3)
            const int tz = b2 - 1 + lane / 9, ty = b1 - 1 + (lane / 3) % 3, tx = b0 - 1 + lane % 3;
            const bool cv = lane < 27 && tz >= 0 && tz < L2 && ty >= 0 && ty < L1 && tx >= 0 && tx < L0;
            const int cc = cv ? (tz * L1 + ty) * L0 + tx : 0;
            int cst = 0, ccnt = 0, gn0 = 0, gn1 = 0, gn2 = 0;
            if (cv) {
                cst = d.cell_start[cc];
                ccnt = d.cell_count[cc];
                gn0 = d.gcount[0][cc];
                gn1 = d.gcount[1][cc];
                gn2 = d.gcount[2][cc];
            }
            const unsigned long long gmask = __ballot(cv && (gn0 + gn1 + gn2) > 0);
            const unsigned long long vmask = __ballot(cv);
            for (int pass = 0; pass < npass; pass++) {
                for (int rz = 0; rz < 3; rz++)
                    for (int ry = 0; ry < 3; ry++) {
                        const int t0 = rz * 9 + ry * 3;
                        const unsigned long long rmask = (vmask >> t0) & 7ull;
                        if (!rmask) continue;
                        const int tlo = t0 + __builtin_ctzll(rmask);
                        const int thi = t0 + 63 - __builtin_clzll(rmask);
                        const bool ghosty = ((gmask >> t0) & 7ull) != 0;
                        // ranges of this row: the whole x-run, or each cell without ghosts
                        for (int t = tlo; t <= thi; t++) {
                            if (ghosty && ((gmask >> t) & 1ull)) continue;
                            const int st = __builtin_amdgcn_readlane(cst, t);
                            const int tend = ghosty ? t : thi;
                            const int en = __builtin_amdgcn_readlane(cst, tend) + __builtin_amdgcn_readlane(ccnt, tend);
                            for (int base = st; base < en; base += 64) {
                                const int s = base + lane;
                                bool acc = false;
                                int j = -1;
                                if (s < en) {
                                    const int2 idm = ci[s];
                                    j = idm.x;
                                    if (j != i) {
                                        const double2 pa = cd[2 * s], pb = cd[2 * s + 1];
                                        double dist_sq = 0.0;
                                        double dx = pa.x - xi; dist_sq = dist_sq + dx * dx;
                                        double dy = pa.y - yi; dist_sq = dist_sq + dy * dy;
                                        double dz = pb.x - zi; dist_sq = dist_sq + dz * dz;
                                        double h_avg = 2.0 * 0.5 * (pb.y + hi);
                                        if (dist_sq < h_avg * h_avg) {
                                            const bool same = idm.y == mi;
                                            if (pass == 0 && same) { acc = true; mind = fmin(mind, dist_sq); }
                                            else if (pass == 1 && !same) { acc = true; mind_cp = fmin(mind_cp, dist_sq); }
                                        }
                                    }
                                }
                                const unsigned long long m = __ballot(acc);
                                const int pos = cnt + __popcll(m & below);
                                if (acc && pos < ncap) out[(long)pos * 64] = j;
                                cnt += __popcll(m);
                            }
                            if (!ghosty) break;
                        }
                    }
            }
            if (d.disctype == 1) {
                if (lane == 0 && cnt < ncap) out[(long)cnt * 64] = i;
                cnt++;
            }
            nreal = cnt;
            if (d.boundary && gmask) {
                for (int pass = 0; pass < npass; pass++) {
                    unsigned long long gm = gmask;
                    while (gm) {
                        const int t = __builtin_ctzll(gm);
                        gm &= gm - 1;
                        const int tcell = __builtin_amdgcn_readlane(cc, t);
                        for (int dd = 2; dd >= 0; dd--) {
                            const int gn = __builtin_amdgcn_readlane(dd == 0 ? gn0 : (dd == 1 ? gn1 : gn2), t);
                            if (gn == 0) continue;
                            const int gs = d.gstart[dd][tcell];
                            for (int base = 0; base < gn; base += 64) {
                                const int q = gn - 1 - (base + lane);
                                bool acc = false;
                                int u = -1;
                                if (q >= 0) {
                                    u = np + gs + q;
                                    double dist_sq = 0.0;
                                    double dx = d.x0[0][u] - xi; dist_sq = dist_sq + dx * dx;
                                    double dy = d.x0[1][u] - yi; dist_sq = dist_sq + dy * dy;
                                    double dz = d.x0[2][u] - zi; dist_sq = dist_sq + dz * dz;
                                    double h_avg = 2.0 * 0.5 * (d.hbuf[cur][u] + hi);
                                    if (dist_sq < h_avg * h_avg) {
                                        const bool same = d.pmat[u] == mi;
                                        if (pass == 0 && same) { acc = true; mind = fmin(mind, dist_sq); }
                                        else if (pass == 1 && !same) { acc = true; mind_cp = fmin(mind_cp, dist_sq); }
                                    }
                                }
                                const unsigned long long m = __ballot(acc);
                                const int pos = cnt + __popcll(m & below);
                                if (acc && pos < ncap) out[(long)pos * 64] = u;
                                cnt += __popcll(m);
                            }
                        }
                    }
                }
            }
            mind = sqrt(fmin(wave_min(mind), wave_min(mind_cp)));
        } else {
            if (d.disctype == 1) {
                if (lane == 0 && cnt < ncap) out[0] = i;
                cnt++;
            }
            nreal = cnt;
        }
        if (lane == 0) {
            d.nreal[i] = nreal;
            d.ntot[i] = cnt;
            d.mindist[i] = mind;
        }
        wmax = max(wmax, cnt);
    }
    if (lane == 0 && wmax > 0) {
        atomicMax(&sc->max_count, wmax);
        if (wmax > ncap) {
            atomicOr(&sc->error_flags, 1);
            atomicExch(&sc->abort, 1);
        }
    }
}

// ------------------------------------------------------------ strain pass
// Pair quantities of one list slot of the strain pass.
struct SPair {
    double xj[3], vj[3], vol, w, dwdx[3], cg[3];
};

__device__ __forceinline__ void strain_pair(const DevState& d, int j, const double xi[3], double hi,
                                            SPair& q) {
    SRec r = load_srec(d, j);   // position: the strain pass's (half step back for real j)
    q.xj[0] = r.x;
    q.xj[1] = r.y;
    q.xj[2] = r.z;
    q.vj[0] = r.vx; q.vj[1] = r.vy; q.vj[2] = r.vz;
    q.vol = r.vol;
    const double havg = 0.5 * (hi + r.h);
    double rad;
    kernel_grad(xi[0] - q.xj[0], xi[1] - q.xj[1], xi[2] - q.xj[2], havg, d.pi, q.dwdx, &rad);
    q.w = kernel_w(rad, havg, d.pi);
}

// velocity gradient of particle i (mcm_oldstrain.f90:34-64 when CORR == 0,
// mcm_corr_strain.f90:37-146 when CORR == 1), returned in every lane
template <int CORR>
__device__ __forceinline__ void strain_wave(const DevState& d, double* buf, int i, int n,
                                            const double xi[3], const double vi[3], double hi, double (&gout)[9]) {
    const int lane = threadIdx.x & 63;
    const int* row = nbr_row(d, d.spos[i]);
    const bool two = n > 64;   // second list slot in use (uniform)
    SPair q[WAVE_SLOTS];
#pragma unroll
    for (int s = 0; s < WAVE_SLOTS; s++) {
        if (s > 0 && !two) break;
        const int k = lane + 64 * s;
        const int j = k < n ? row[(long)k * 64] : i;
        strain_pair(d, j, xi, hi, q[s]);
    }
    if (CORR == 0) {
        double t[WAVE_SLOTS][9];
#pragma unroll
        for (int s = 0; s < WAVE_SLOTS; s++) {
            if (s > 0 && !two) break;
            const double dv[3] = {q[s].vj[0] - vi[0], q[s].vj[1] - vi[1], q[s].vj[2] - vi[2]};
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++) t[s][3 * a + b] = q[s].vol * dv[a] * q[s].dwdx[b];
        }
        ordered_sums<9>(buf, t, n, gout);
        return;
    }
    // pass 1: Shepard sum and gradient sum
    double sums[4];
    {
        double t[WAVE_SLOTS][4];
#pragma unroll
        for (int s = 0; s < WAVE_SLOTS; s++) {
            if (s > 0 && !two) break;
            t[s][0] = q[s].vol * q[s].w;
#pragma unroll
            for (int l = 0; l < 3; l++) t[s][1 + l] = q[s].vol * q[s].dwdx[l];
        }
        ordered_sums<4>(buf, t, n, sums);
    }
    const double sw = sums[0], sg[3] = {sums[1], sums[2], sums[3]};
    const double sw2 = sw * sw;
    const Divisor isw2(sw2);
    // pass 2: B(n,m) = sum V (xj - xi)_n cg_m
    double B[9], invB[9];
    {
        double t[WAVE_SLOTS][9];
#pragma unroll
        for (int s = 0; s < WAVE_SLOTS; s++) {
            if (s > 0 && !two) break;
            isw2.div3(q[s].dwdx[0] * sw - q[s].w * sg[0], q[s].dwdx[1] * sw - q[s].w * sg[1], q[s].dwdx[2] * sw - q[s].w * sg[2], q[s].cg);
#pragma unroll
            for (int m = 0; m < 3; m++)
#pragma unroll
                for (int nn = 0; nn < 3; nn++) t[s][nn + 3 * m] = q[s].vol * (q[s].xj[nn] - xi[nn]) * q[s].cg[m];
        }
        ordered_sums<9>(buf, t, n, B);
    }
    inverse3(B, invB);
    // pass 3: L(n,m) = sum V vj_n (invB cg)_m
    {
        double t[WAVE_SLOTS][9];
#pragma unroll
        for (int s = 0; s < WAVE_SLOTS; s++) {
            if (s > 0 && !two) break;
            double gg[3];
#pragma unroll
            for (int r = 0; r < 3; r++)
                gg[r] = invB[r] * q[s].cg[0] + invB[r + 3] * q[s].cg[1] + invB[r + 6] * q[s].cg[2];
#pragma unroll
            for (int nn = 0; nn < 3; nn++)
#pragma unroll
                for (int m = 0; m < 3; m++) t[s][3 * nn + m] = q[s].vol * q[s].vj[nn] * gg[m];
        }
        ordered_sums<9>(buf, t, n, gout);
    }
}

// Strain pass of the fused step, wavefront per particle: stages rod, spin
// and tracerod of every particle, x after the pass (x1) and the momentum
// record's position.  k_strain_const<MODE_FCONST> then runs update_h, the
// constitutive update and the dt reduction thread per particle, so that its
// register demand does not limit the occupancy of the gathers.
template <int CORR>
__global__ __launch_bounds__(256) void k_strain_wave(DevState d_, int cur) {
    __shared__ __align__(16) double lds[4][OSUM_WORDS];
    Blk bk;
    const DevState& d = enter<true>(d_, bk);
    const int lb = bk.b;
    const int lane = threadIdx.x & 63;
    const int wid = wave_uniform(threadIdx.x >> 6);
    double* buf = lds[wid];
    Scalars* sc = d.sc;
    if (sc->abort) return;
    const double dt = sc->dt;
    const int wbase = lb * 256 + wid * 64;
    double mine[9];
#pragma unroll
    for (int c = 0; c < 9; c++) mine[c] = 0.0;
    for (int p = 0; p < 64; p++) {
        const int i = wbase + p;
        if (i >= d.np) break;
        const int n = d.ntot[i];
        double xi[3], vi[3];
        for (int a = 0; a < 3; a++) {
            vi[a] = d.v[a][i];
            xi[a] = d.x0[a][i] - 0.5 * vi[a] * dt;
        }
        const double hi = d.hbuf[cur][i];
        double gsum[9];
        strain_wave<CORR>(d, buf, i, n, xi, vi, hi, gsum);
        keep_if_lane<9>(p, gsum, mine);
    }
    const int i = wbase + lane;
    if (i < d.np) {
        PState s;
        double g[3][3];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) g[a][b] = mine[3 * a + b];
        rod_spin(g, s);
        for (int k = 0; k < 6; k++) d.rod[k][i] = s.rod[k];
        for (int k = 0; k < 3; k++) d.spin[k][i] = s.spin[k];
        d.tracerod[i] = s.tracerod;
        double xi[3], vi[3];
        for (int a = 0; a < 3; a++) {
            vi[a] = d.v[a][i];
            xi[a] = d.x0[a][i] - 0.5 * vi[a] * dt;
        }
        strain_finish(d, i, dt, xi, vi);
    }
}

// ---------------------------------------------------------- momentum pass
struct MPair {
    double dx[3], vol, w, dwdx[3], A[6], m, cg[3];
};

// acceleration of particle i (mcm_oldacceleration.f90:60-89 when CORR == 0,
// mcm_corr_acceleration.f90:37-163 when CORR == 1), returned in every lane
template <int CORR>
__device__ __forceinline__ void accel_wave(const DevState& d, double* buf, int i, int n, int self_k, const MRec& me,
                                           double (&acc)[3]) {
    const int lane = threadIdx.x & 63;
    const int* row = nbr_row(d, d.spos[i]);
    const bool two = n > 64;   // second list slot in use (uniform)
    MPair q[WAVE_SLOTS];
#pragma unroll
    for (int s = 0; s < WAVE_SLOTS; s++) {
        if (s > 0 && !two) break;
        const int k = lane + 64 * s;
        const int j = k < n ? row[(long)k * 64] : i;
        MRec r = load_mrec(d, j);
        const double havg = 0.5 * (me.h + r.h);
        double rad;
        if (CORR == 0) {
            kernel_grad(me.x - r.x, me.y - r.y, me.z - r.z, havg, d.pi, q[s].dwdx, &rad);
            q[s].w = 0.0;
        } else {
            // kernel centred at the neighbour (mcm_corr_acceleration.f90:73-74)
            kernel_grad(r.x - me.x, r.y - me.y, r.z - me.z, havg, d.pi, q[s].dwdx, &rad);
            q[s].w = kernel_w(rad, havg, d.pi);
        }
        q[s].dx[0] = r.x - me.x; q[s].dx[1] = r.y - me.y; q[s].dx[2] = r.z - me.z;
        q[s].vol = r.vol;
        q[s].m = r.m;
#pragma unroll
        for (int c = 0; c < 6; c++) q[s].A[c] = r.A[c];
    }
    if (CORR == 0) {
        double t[WAVE_SLOTS][3];
#pragma unroll
        for (int s = 0; s < WAVE_SLOTS; s++) {
            if (s > 0 && !two) break;
#pragma unroll
            for (int nn = 0; nn < 3; nn++) {
                double ds = 0.0;
#pragma unroll
                for (int mm = 0; mm < 3; mm++) {
                    const int c = sym_index(mm, nn);
                    ds = ds + (q[s].A[c] + me.A[c]) * q[s].dwdx[mm];
                }
                t[s][nn] = q[s].m * ds;
            }
        }
        ordered_sums<3>(buf, t, n, acc);
        return;
    }
    // mass of the last list entry (mcm_corr_acceleration.f90:155-161)
    const int kl = n - 1;
    const double massj = kl < 64 ? lane_value(q[0].m, kl) : lane_value(q[1].m, kl - 64);
    double sums[4];
    {
        double t[WAVE_SLOTS][4];
#pragma unroll
        for (int s = 0; s < WAVE_SLOTS; s++) {
            if (s > 0 && !two) break;
            t[s][0] = q[s].vol * q[s].w;
#pragma unroll
            for (int l = 0; l < 3; l++) t[s][1 + l] = q[s].vol * q[s].dwdx[l];
        }
        ordered_sums<4>(buf, t, n, sums);
    }
    const double sw = sums[0], sg[3] = {sums[1], sums[2], sums[3]};
    const double sw2 = sw * sw;
    const Divisor isw2(sw2);
    double B[9], invB[9];
    {
        double t[WAVE_SLOTS][9];
#pragma unroll
        for (int s = 0; s < WAVE_SLOTS; s++) {
            if (s > 0 && !two) break;
            isw2.div3(q[s].dwdx[0] * sw - q[s].w * sg[0], q[s].dwdx[1] * sw - q[s].w * sg[1], q[s].dwdx[2] * sw - q[s].w * sg[2], q[s].cg);
#pragma unroll
            for (int m = 0; m < 3; m++)
#pragma unroll
                for (int nn = 0; nn < 3; nn++) t[s][nn + 3 * m] = q[s].vol * q[s].dx[nn] * q[s].cg[m];
        }
        ordered_sums<9>(buf, t, n, B);
    }
    inverse3(B, invB);
    // deltasig(c, l) = sum over the list but the self slot of dsig_c cg_l, in
    // two rounds of three stress components; the self slot adds -0.0, which
    // leaves every sum unchanged
    double D[6][3];
#pragma unroll
    for (int half = 0; half < 2; half++) {
        double t[WAVE_SLOTS][9], r9[9];
#pragma unroll
        for (int s = 0; s < WAVE_SLOTS; s++) {
            if (s > 0 && !two) break;
            const bool self = lane + 64 * s == self_k;
#pragma unroll
            for (int cc = 0; cc < 3; cc++) {
                const int c = 3 * half + cc;
                const double dsig = q[s].A[c] + me.A[c];
#pragma unroll
                for (int l = 0; l < 3; l++) t[s][3 * cc + l] = self ? -0.0 : dsig * q[s].cg[l];
            }
        }
        ordered_sums<9>(buf, t, n, r9);
#pragma unroll
        for (int cc = 0; cc < 3; cc++)
#pragma unroll
            for (int l = 0; l < 3; l++) D[3 * half + cc][l] = r9[3 * cc + l];
    }
#pragma unroll
    for (int nn = 0; nn < 3; nn++) {
        double a = 0.0;
#pragma unroll
        for (int mm = 0; mm < 3; mm++) {
            const int c = sym_index(nn, mm);
#pragma unroll
            for (int l = 0; l < 3; l++) a = a + massj * D[c][l] * invB[l + 3 * mm];
        }
        acc[nn] = a;
    }
}

// Fused momentum + boundcond + energies + velocity + move; the last
// workgroup reduces energies and coordinate bounds.
template <int CORR>
__global__ __launch_bounds__(256) void k_momentum_wave(DevState d_, int cur) {
    __shared__ __align__(16) double lds[4][OSUM_WORDS];
    Blk bk;
    const DevState& d = enter<true>(d_, bk);
    const int lb = bk.b;
    const int lane = threadIdx.x & 63;
    const int wid = wave_uniform(threadIdx.x >> 6);
    double* buf = lds[wid];
    Scalars* sc = d.sc;
    const bool abort = sc->abort;
    const int wbase = lb * 256 + wid * 64;
    double mine[3] = {0.0, 0.0, 0.0};
    if (!abort) {
        for (int p = 0; p < 64; p++) {
            const int i = wbase + p;
            if (i >= d.np) break;
            const int n = d.ntot[i];
            const int self_k = d.nreal[i] - 1;
            const MRec me = load_mrec(d, i);
            double acc[3];
            accel_wave<CORR>(d, buf, i, n, self_k, me, acc);
            keep_if_lane<3>(p, acc, mine);
        }
    }
    CPart cp;
    init_c_part(cp);
    const int i = wbase + lane;
    if (!abort && i < d.np) {
        const MRec me = load_mrec(d, i);
        const double xi[3] = {me.x, me.y, me.z};
        double acc[3] = {mine[0], mine[1], mine[2]};
        integrate_particle<false, true>(d, i, acc, xi, cp);
    }
    reduce_bound_partials(d, cp, true, lb);   // energies: the constitutive pass
}
