// mcm_nbr.hip -- neighbour search of one MCM step on gfx950.
//
// Replaces mcm_setuplist.f90:30-135, mcm_ghost_setup.f90:32-425 and
// mcm_ll_neighbours.f90:44-311.  The reference pushes particles onto
// per-cell linked lists every step; here the cells are built by a counting
// sort (atomic histogram, exclusive scan, scatter) into cell-sorted copies of
// x and h, and each cell's members are ordered by descending particle id, the
// order the reference's pushed lists are walked in.  Ghost (symmetry /
// periodic) particles are generated on the device, plane by plane, with ids
// in the reference's creation order, so the neighbour lists come out in the
// reference's order: cells in z-y-x order, descending id within a cell,
// same-material before other-material, self (corrected SPH), then ghosts.
#pragma once
#include "mcm_device.h"

// ---------------------------------------------------------------- helpers
// min of two doubles that are not NaN, in one instruction (fmin also
// quiets signalling NaNs, two more instructions)
__device__ __forceinline__ double min_f64(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

__device__ __forceinline__ int wave_incl_scan(int x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

// exclusive scan over a 1024-thread block; *total receives the block sum
__device__ int block_excl_scan_1024(int v, int* total) {
    __shared__ int wsum[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int inc = wave_incl_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        int w = lane < 16 ? wsum[lane] : 0;
        w = wave_incl_scan(w);
        if (lane < 16) wsum[lane] = w;
    }
    __syncthreads();
    int base = wid > 0 ? wsum[wid - 1] : 0;
    *total = wsum[15];
    __syncthreads();
    return base + inc - v;
}

// (int)(t) + 1 as the reference computes a 1-based cell index; out-of-range
// values map to 0 (outside the sort domain)
__device__ __forceinline__ int cell_index(double t) {
    if (!(t > -1.0e9 && t < 1.0e9)) return 0;
    return (int)t + 1;
}

// ------------------------------------------------------------ grid set-up
// mcm_setuplist.f90:240-291, on the device from hmax and the last coordinate
// bounds; also resets the per-step counters.
__global__ void k_grid_setup(DevState d_) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    Scalars* sc = d.sc;
    d.range[0] = 0;   // the search rewrites every pair record; its writers flag them again
    sc->nout = 0;
    GridDev g;
    const double factor = 2.0;
    for (int i = 0; i < 3; i++) {
        g.gridmin[i] = 0.0;
        g.lim[i] = 1;
        g.gridsize[i] = 1.0 / (factor * sc->hmax);
    }
    bool bad = false;
    for (int i = 0; i < 3; i++) {
        double t = 0.0;
        switch (d.btype[i]) {
            case 1:
                g.gridmin[i] = sc->coord_min[i] - 0.01 * sc->hmax;
                t = g.gridsize[i] * (sc->coord_max[i] - g.gridmin[i]);
                break;
            case 2:
                g.gridmin[i] = d.bx[i][0];
                t = g.gridsize[i] * (sc->coord_max[i] - g.gridmin[i]);
                break;
            case 3:
                t = g.gridsize[i] * (d.bx[i][1] - sc->coord_min[i]);
                break;
            case 4:
                t = g.gridsize[i] * (d.bx[i][1] - d.bx[i][0]) - 1.0;
                break;
        }
        if (!(t > -1.0e9 && t < 1.0e9)) { bad = true; t = 0.0; }
        g.lim[i] = (int)t + 1;
        if (d.btype[i] == 3) g.gridmin[i] = d.bx[i][1] - g.lim[i] * factor * sc->hmax;
        if (d.btype[i] == 4) {
            if (g.lim[i] < 1) g.lim[i] = 1;
            g.gridsize[i] = (double)(float)g.lim[i] / (d.bx[i][1] - d.bx[i][0]);
        }
    }
    for (int i = 0; i < 3; i++) {
        int c0 = d.bcode[i][0], c1 = d.bcode[i][1];
        if (c0 == 1 || c0 == 2) {
            g.lim[i] += 1;
            g.gridmin[i] = g.gridmin[i] - (1.0 / g.gridsize[i]);
        }
        if (c1 == 1 || c1 == 2) g.lim[i] += 1;
        if (g.lim[i] < 1) bad = true;
    }
    long long ncell = (long long)g.lim[0] * g.lim[1] * g.lim[2];
    if (sc->abort) return;   // sticky until the host has grown the capacity
    if (bad || ncell > (long long)d.cellcap) {
        sc->error_flags |= 4;
        sc->abort = 1;
        g.ncell = 0;
    } else {
        g.ncell = (int)ncell;
    }
    sc->grid = g;
    sc->ngp = 0;
    sc->max_count = 0;
    sc->hcap_bits = 0;
    for (int k = 0; k < 3; k++) {
        sc->skin_dmin[k] = ~0ull;
        sc->skin_dmax[k] = 0ull;
    }
    sc->skin_dh = 0ull;
    sc->skin_near = 0;
    sc->skin_rebuild = 0;
}

__global__ void k_clear_cells(DevState d_) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    int ncell = d.sc->grid.ncell;
    for (int c = bk.b * blockDim.x + threadIdx.x; c < ncell; c += bk.n * blockDim.x) {
        d.cell_count[c] = 0;
        d.gcount[0][c] = 0; d.gcount[1][c] = 0; d.gcount[2][c] = 0;
        d.gany[c] = 0;
    }
}

// The strain records' half-step positions of the real particles with the
// current dt.  mcm_initial sets the first dt after its neighbour search, so
// it refreshes them; a strain pass on those lists then shifts by that dt.
__global__ __launch_bounds__(MCM_BLOCK) void k_rs_shift(DevState d) {
    const int i = blockIdx.x * MCM_BLOCK + threadIdx.x;
    if (i >= d.np) return;
    const double dt = d.sc->dt;
    const double vx = d.v[0][i], vy = d.v[1][i], vz = d.v[2][i];
    const double s0 = d.x0[0][i] - 0.5 * vx * dt, s1 = d.x0[1][i] - 0.5 * vy * dt, s2 = d.x0[2][i] - 0.5 * vz * dt;
    *rs_at(d, 0, i) = make_double2(s0, s1);
    *lo_of(rs_at(d, 1, i)) = s2;
    if (!(fast_coord(s0) & fast_coord(s1) & fast_coord(s2))) atomicOr(d.range, (int)RANGE_RS);
}

// ------------------------------------------------------- neighbour skin
// With a skin the real partners of a particle are taken from its skin list:
// every real particle j with |x_i - x_j| < h_i + h_j + w at the last build
// (w the skin width), in the walk order of that build.  Each step keeps
// those with the reference's test at the current x and h, so the lists hold
// the reference's sets (mcm_ll_neighbours.f90:140-148), not its order.
//
// Why the skin holds every partner.  At the build the particles are put in
// cubic blocks of edge b >= 2 (2 hc + w) (hc the largest h then), by their
// position x_b then.  Let D_k = x_k - x_b,k be the displacement of particle
// k since the build and E_k = |h_k - h_b,k|.  Take a pair (i, j) that is not
// in the skin: |x_b,i - x_b,j| >= h_b,i + h_b,j + w.
//  * i and j in the same or adjacent blocks A, B: with R_AB the diagonal of
//    the box spanned by the D_k of both blocks and H_AB the largest E_k there,
//      |x_i - x_j| >= |x_b,i - x_b,j| - |D_i - D_j|
//                 >= h_b,i + h_b,j + w - R_AB >= h_i + h_j + w - R_AB - 2 H_AB,
//    so the pair fails the reference's test while R_AB + 2 H_AB < w;
//  * blocks further apart: more than b separates x_b,i and x_b,j along some
//    axis, so with R and H the same extents over all particles
//      |x_i - x_j| >= |x_b,i - x_b,j| - |D_i - D_j| > b - R,
//    while h_i + h_j <= 2 (hc + H): the pair fails the test while
//    R + 2 H <= b - 2 hc.
// Every step records the per-block extents of D and E (k_skin_stats) and
// checks both conditions (k_skin_blocks, k_skin_check); the lists are built
// again in the first step that breaks one of them.  The build takes w wider
// by a relative 1e-9, and the checks compare with a margin of the same size,
// far above the rounding of the distances.
//
// The skin lists hold sorted positions of the build, and the candidates
// are read from xh by that position, {x, y, z, h} of this step: particles
// that were neighbours in space at the build lie near each other in that
// order, so the gathers of a wavefront share cache lines as the list
// build's candidate records do.  h is negated for a particle that is no
// candidate this step (outside the grid, or in a cell holding ghosts, which
// hides its real particles as in the reference's linked lists).

// xh of real particle i with valid skin lists (binned: it lies in the grid)
__device__ __forceinline__ void skin_write_xh(const DevState& d, int i, bool binned, int cur) {
    double2* xr = reinterpret_cast<double2*>(d.xh) + 2 * (long)d.skin_pos[i];
    const double h = d.hbuf[cur][i];
    xr[0] = make_double2(d.x0[0][i], d.x0[1][i]);
    xr[1] = make_double2(d.x0[2][i], binned ? h : -h);
}

// Per-block extents of the displacement since the build: one thread per
// sorted position of the build; the lanes of a wavefront mostly share a
// block, so each distinct block of the wavefront takes one atomic per
// value.  Buffer skin_bstat[cur] (cleared in the step before).
__global__ __launch_bounds__(MCM_BLOCK) void k_skin_stats(DevState d_, int cur) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    Scalars* sc = d.sc;
    if (sc->abort || !sc->skin_valid) return;
    const int tb = bk.b * blockDim.x + threadIdx.x;
    int blk = -1;
    double dv[3] = {0.0, 0.0, 0.0}, e = 0.0;
    (void)cur;
    if (tb < sc->skin_nbuilt) {
        const double2* br = reinterpret_cast<const double2*>(d.skin_base) + 2 * (long)tb;
        const double2* xr = reinterpret_cast<const double2*>(d.xh) + 2 * (long)tb;
        const double2 b0 = br[0], b1 = br[1], x0 = xr[0], x1 = xr[1];
        dv[0] = x0.x - b0.x;
        dv[1] = x0.y - b0.y;
        dv[2] = x1.x - b1.x;
        e = fabs(fabs(x1.y) - b1.y);
        // a NaN position or h makes the skin useless: force a build
        if (!(dv[0] == dv[0] && dv[1] == dv[1] && dv[2] == dv[2] && e == e)) e = 1.0e300;
        blk = d.skin_blk[tb];
    }
    unsigned long long* st = d.skin_bstat[cur];
    unsigned long long todo = __ballot(blk >= 0);
    while (todo) {
        const int leader = __builtin_ctzll(todo);
        const int lb = __builtin_amdgcn_readlane(blk, leader);
        const bool mine = blk == lb;
        double lo[3], hi[3], em;
        for (int k = 0; k < 3; k++) {
            lo[k] = wave_min(mine ? dv[k] : 1.0e300);
            hi[k] = wave_max(mine ? dv[k] : -1.0e300);
        }
        em = wave_max(mine ? e : 0.0);
        if ((threadIdx.x & 63) == leader) {
            unsigned long long* b = st + 8 * (long)lb;
            for (int k = 0; k < 3; k++) {
                atomicMin(&b[k], ordbits(lo[k]));
                atomicMax(&b[3 + k], ordbits(hi[k]));
            }
            atomicMax(&b[6], ordbits(em));
        }
        todo &= ~__ballot(mine);
    }
}

// One thread per block A of the build: the joint extent of A with each
// adjacent block B (B after A in linear order, and A itself) against the
// skin width; folds the extents of all blocks into the run's totals; clears
// the other buffer for the next step.
__global__ __launch_bounds__(MCM_BLOCK) void k_skin_blocks(DevState d_, int cur) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    Scalars* sc = d.sc;
    if (sc->abort) return;
    unsigned long long* clr = d.skin_bstat[cur ^ 1];
    for (long q = (long)bk.b * blockDim.x + threadIdx.x; q < 8L * d.skin_blkcap; q += (long)bk.n * blockDim.x)
        clr[q] = (q & 7) < 3 ? ~0ull : 0ull;
    if (!sc->skin_valid) return;
    const unsigned long long* st = d.skin_bstat[cur];
    const int n0 = sc->skin_bdim[0], n1 = sc->skin_bdim[1], n2 = sc->skin_bdim[2];
    const int nblk = n0 * n1 * n2;
    const double w = sc->skin_len;
    double lo[3] = {1.0e300, 1.0e300, 1.0e300}, hi[3] = {-1.0e300, -1.0e300, -1.0e300}, em = 0.0;
    int near = 0;
    for (int a = bk.b * blockDim.x + threadIdx.x; a < nblk; a += bk.n * blockDim.x) {
        const unsigned long long* A = st + 8 * (long)a;
        if (A[0] > A[3]) continue;   // no particle in the block
        double alo[3], ahi[3];
        for (int k = 0; k < 3; k++) {
            alo[k] = ordval(A[k]);
            ahi[k] = ordval(A[3 + k]);
            lo[k] = fmin(lo[k], alo[k]);
            hi[k] = fmax(hi[k], ahi[k]);
        }
        const double ae = ordval(A[6]);
        em = fmax(em, ae);
        const int a0 = a % n0, a1 = (a / n0) % n1, a2 = a / (n0 * n1);
        for (int o = 13; o < 27; o++) {   // offsets (dz, dy, dx) with B >= A
            const int c0 = a0 + o % 3 - 1, c1 = a1 + (o / 3) % 3 - 1, c2 = a2 + o / 9 - 1;
            if (c0 < 0 || c0 >= n0 || c1 < 0 || c1 >= n1 || c2 < 0 || c2 >= n2) continue;
            const unsigned long long* B = st + 8 * (long)((c2 * n1 + c1) * n0 + c0);
            if (B[0] > B[3]) continue;
            double r2 = 0.0;
            for (int k = 0; k < 3; k++) {
                const double ext = fmax(ahi[k], ordval(B[3 + k])) - fmin(alo[k], ordval(B[k]));
                r2 += ext * ext;
            }
            const double bound = (sqrt(r2) + 2.0 * fmax(ae, ordval(B[6]))) * (1.0 + 1.0e-9);
            if (!(bound < w)) near = 1;
        }
    }
    // the workgroup's extents and flag, folded by k_skin_check
    __shared__ double red[8][MCM_BLOCK / 64];
    for (int k = 0; k < 3; k++) { lo[k] = wave_min(lo[k]); hi[k] = wave_max(hi[k]); }
    em = wave_max(em);
    const double fl = __any(near) ? 1.0 : 0.0;
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        for (int k = 0; k < 3; k++) { red[k][wv] = lo[k]; red[3 + k][wv] = hi[k]; }
        red[6][wv] = em;
        red[7][wv] = fl;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        const int k = threadIdx.x;
        double v = red[k][0];
        for (int q = 1; q < MCM_BLOCK / 64; q++) v = k < 3 ? fmin(v, red[k][q]) : fmax(v, red[k][q]);
        d.skin_red[8 * (long)bk.b + k] = v;
    }
}

// One thread: whether the skin lists must be built in this step, and if so
// their width and block grid.  A build is needed when they are not valid
// (first search, grown capacities, replaced state), when one of the two
// conditions above fails, or when a particle lay outside the grid at the
// last build (it is in no skin list, and may be back in the grid now).
__global__ __launch_bounds__(64) void k_skin_check(DevState d_) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    Scalars* sc = d.sc;
    if (sc->abort) return;
    // fold the partial extents of k_skin_blocks (one wavefront)
    double lo[3] = {1.0e300, 1.0e300, 1.0e300}, hi[3] = {-1.0e300, -1.0e300, -1.0e300}, em = 0.0, fl = 0.0;
    if (sc->skin_valid) {
        for (int q = threadIdx.x; q < d.skin_nred; q += 64) {
            const double* r = d.skin_red + 8 * (long)q;
            for (int k = 0; k < 3; k++) { lo[k] = fmin(lo[k], r[k]); hi[k] = fmax(hi[k], r[3 + k]); }
            em = fmax(em, r[6]);
            fl = fmax(fl, r[7]);
        }
        for (int k = 0; k < 3; k++) { lo[k] = wave_min(lo[k]); hi[k] = wave_max(hi[k]); }
        em = wave_max(em);
        fl = wave_max(fl);
    }
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 3; k++) { sc->skin_dmin[k] = ordbits(lo[k]); sc->skin_dmax[k] = ordbits(hi[k]); }
    sc->skin_dh = ordbits(em);
    sc->skin_near = fl > 0.0 ? 1 : 0;
    bool need = !sc->skin_valid || fl > 0.0;
    if (!need) {
        double r2 = 0.0;
        for (int k = 0; k < 3; k++) {
            const double e = hi[k] - lo[k];
            r2 += e * e;
        }
        const double bound = (sqrt(r2) + 2.0 * em) * (1.0 + 1.0e-9);
        need = !(bound < sc->skin_bedge - 2.0 * sc->skin_hcap);
    }
    sc->skin_rebuild = need ? 1 : 0;
    if (!need) return;
    const double hc = __longlong_as_double((long long)sc->hcap_bits);
    const double w = d.skin * hc;
    sc->skin_len = w;
    sc->skin_hcap = hc;
    sc->skin_valid = sc->nout == 0 ? 1 : 0;
    sc->skin_builds += 1;
    // blocks over the particles' bounds (all lie inside the search grid,
    // which starts at gridmin); the edge doubles until the blocks fit
    double edge = 2.0 * (2.0 * hc + w) * (1.0 + 1.0e-9);
    int n[3];
    for (;;) {
        long long tot = 1;
        for (int k = 0; k < 3; k++) {
            const double span = (sc->coord_max[k] - sc->grid.gridmin[k]) / edge;
            n[k] = span > 0.0 && span < 1.0e9 ? (int)span + 1 : 1;
            tot *= n[k];
        }
        if (tot <= d.skin_blkcap) break;
        edge *= 2.0;
    }
    for (int k = 0; k < 3; k++) {
        sc->skin_borig[k] = sc->grid.gridmin[k];
        sc->skin_bdim[k] = n[k];
    }
    sc->skin_bedge = edge;
}

// histogram: cell of every real particle (mcm_setuplist.f90:320-345); also
// lays down the particle's strain gather record for this step, whose position
// is already the half-step position x - v dt / 2 of the strain pass
// (mcm_strain.f90:27-31) with this step's dt, the expression and rounding of
// the pair loops' own shift
__global__ __launch_bounds__(256) void k_bin(DevState d_, int cur) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    const int i = bk.b * blockDim.x + threadIdx.x;
    if (d.sc->abort) return;
    const bool live = i < d.np;
    double hb = 0.0;   // h of a binned particle: bounds every list partner's h
    int cell = -1;
    if (live) {
        const GridDev g = d.sc->grid;
        const double x[3] = {d.x0[0][i], d.x0[1][i], d.x0[2][i]};
        const double h = d.hbuf[cur][i];
        {
            const double dt = d.sc->dt;
            double vx = d.v[0][i], vy = d.v[1][i], vz = d.v[2][i];
            const double s0 = x[0] - 0.5 * vx * dt, s1 = x[1] - 0.5 * vy * dt, s2 = x[2] - 0.5 * vz * dt;
            *rs_at(d, 0, i) = make_double2(s0, s1);
            *rs_at(d, 1, i) = make_double2(s2, h);
            *rs_at(d, 2, i) = make_double2(d.mass[i] / d.rbuf[cur][i], vx);
            *rs_at(d, 3, i) = make_double2(vy, vz);
            if (!(fast_coord(s0) & fast_coord(s1) & fast_coord(s2) & fast_h(h))) atomicOr(d.range, (int)RANGE_RS);
        }
        int box[3];
        bool in = true;
        for (int j = 0; j < 3; j++) {
            box[j] = cell_index(g.gridsize[j] * (x[j] - g.gridmin[j]));
            if (box[j] < 1 || box[j] > g.lim[j]) in = false;
        }
        if (in) {
            cell = ((box[2] - 1) * g.lim[1] + (box[1] - 1)) * g.lim[0] + (box[0] - 1);
            hb = h;
        } else {
            // outside the grid: no neighbours, only itself in corrected SPH
            // (mcm_ll_neighbours.f90:109, 288-299); the last positions of the
            // sorted order
            const int t = d.np - 1 - atomicAdd(&d.sc->nout, 1);
            d.sorted[t] = i;
            d.spos[i] = t;
            int cnt = 0;
            if (d.disctype == 1) {
                nbr_row(d, t)[0] = i;
                cnt = 1;
            }
            d.nreal[i] = cnt;
            d.ntot[i] = cnt;
            d.mindist[i] = 1.0e20;
            if (d.contacttype > 0) d.ncont[i] = 0;
        }
        d.cell_of[i] = cell;
    }
    if (d.skin > 0.0 && live && d.sc->skin_valid) skin_write_xh(d, i, cell >= 0, cur);
    // histogram with one atomic per run of lanes in the same cell, all runs'
    // atomics issued together: the lanes of a run take consecutive slots (the
    // cell sort orders every cell's members by id later, so any distinct slots
    // do).  Particles in id order sit in runs along the lattice rows.
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(cell, 1, 64);
    const bool head = cell >= 0 && (lane == 0 || prev != cell);
    const unsigned long long heads = __ballot(head);
    const unsigned long long bounds = __ballot(head || cell < 0);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);   // lanes 0 .. lane
    const unsigned long long after = bounds & ~upto;
    const int next = after ? __builtin_ctzll(after) : 64;   // first lane past this lane's run
    int base = 0;
    if (head) base = atomicAdd(&d.cell_count[cell], next - lane);
    const unsigned long long mine = heads & upto;
    const int hl = mine ? 63 - __builtin_clzll(mine) : lane;   // the run's head
    base = __shfl(base, hl, 64);
    if (cell >= 0) d.slot_of[i] = base + (lane - hl);
    // workgroup maximum, one atomic per workgroup (positive doubles order as
    // their bit patterns)
    __shared__ double wmax[4];
    hb = wave_max(hb);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = hb;
    __syncthreads();
    if (threadIdx.x == 0) {
        hb = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
        if (hb > 0.0) atomicMax(&d.sc->hcap_bits, (unsigned long long)__double_as_longlong(hb));
    }
}

// exclusive scan of cell_count -> cell_start, in three launches; the tile
// kernels take tiles b, b + n, ... of the grid's cells
#define SCAN_ITEMS 4
#define SCAN_TILE (1024 * SCAN_ITEMS)
__global__ __launch_bounds__(1024) void k_scan_tiles(DevState d_) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    const int ncell = d.sc->grid.ncell;
    for (int b = bk.b; b * SCAN_TILE < ncell; b += bk.n) {
        const int base = b * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
        int s = 0;
        for (int k = 0; k < SCAN_ITEMS; k++) s += (base + k < ncell) ? d.cell_count[base + k] : 0;
        int total;
        block_excl_scan_1024(s, &total);
        if (threadIdx.x == 0) d.scan_tmp[b] = total;
    }
}

__global__ __launch_bounds__(1024) void k_scan_top(DevState d_) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    int ntile = (d.sc->grid.ncell + SCAN_TILE - 1) / SCAN_TILE;
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int c0 = 0; c0 < ntile; c0 += 1024) {
        int k = c0 + threadIdx.x;
        int v = k < ntile ? d.scan_tmp[k] : 0;
        int total;
        int ex = block_excl_scan_1024(v, &total);
        if (k < ntile) d.scan_tmp[k] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void k_scan_apply(DevState d_) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    const int ncell = d.sc->grid.ncell;
    for (int b = bk.b; b * SCAN_TILE < ncell; b += bk.n) {
        const int base = b * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
        int vals[SCAN_ITEMS];
        int s = 0;
        for (int k = 0; k < SCAN_ITEMS; k++) {
            vals[k] = (base + k < ncell) ? d.cell_count[base + k] : 0;
            s += vals[k];
        }
        int total;
        int ex = block_excl_scan_1024(s, &total) + d.scan_tmp[b];
        for (int k = 0; k < SCAN_ITEMS; k++) {
            if (base + k < ncell) d.cell_start[base + k] = ex;
            ex += vals[k];
        }
    }
}

// each binned particle into its cell's run, in histogram slot order
__global__ __launch_bounds__(256) void k_scatter(DevState d_) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    int i = bk.b * blockDim.x + threadIdx.x;
    if (i >= d.np || d.sc->abort) return;
    int c = d.cell_of[i];
    if (c >= 0) d.unsorted[d.cell_start[c] + d.slot_of[i]] = i;
}

// order each cell by descending id (the walk order of the reference's pushed
// lists), one thread per particle: its place in the cell's run is the number
// of larger ids in the cell (ids are distinct)
__global__ __launch_bounds__(256) void k_sort_cells(DevState d_) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    const int i = bk.b * blockDim.x + threadIdx.x;
    if (i >= d.np || d.sc->abort) return;
    const int c = d.cell_of[i];
    if (c < 0) return;
    const int s0 = d.cell_start[c], n = d.cell_count[c];
    const int* run = d.unsorted + s0;
    int rank = 0;
    for (int k = 0; k < n; k++) rank += run[k] > i ? 1 : 0;
    d.sorted[s0 + rank] = i;
}

// cell-sorted copies of x, h and the material of the binned particles, one
// thread per sorted slot (the candidate records of the list build)
__global__ __launch_bounds__(256) void k_fill_cand(DevState d_, int cur) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    const int s = bk.b * blockDim.x + threadIdx.x;
    const GridDev& g = d.sc->grid;
    if (d.sc->abort || g.ncell <= 0) return;
    const int nsorted = d.cell_start[g.ncell - 1] + d.cell_count[g.ncell - 1];
    if (s >= nsorted) return;
    const int j = d.sorted[s];
    d.spos[j] = s;
    double2* cd = reinterpret_cast<double2*>(d.cand) + 2 * (long)s;
    cd[0] = make_double2(d.x0[0][j], d.x0[1][j]);
    cd[1] = make_double2(d.x0[2][j], d.hbuf[cur][j]);
    reinterpret_cast<int2*>(d.cand_id)[s] = make_int2(j, d.pmat[j]);
}

// ------------------------------------------------------------------ ghosts
__device__ __forceinline__ void make_ghost(const DevState& d, int cur, const PlaneDev& P, bool from_real, int src,
                                           int gid) {
    int u = d.np + gid;
    int us = from_real ? src : d.np + src;
    for (int k = 0; k < 3; k++) {
        double xv = d.x0[k][us] * P.mult[k] + P.add[k];
        d.x0[k][u] = xv;
        d.x1[k][u] = xv;
        d.v[k][u] = d.v[k][us] * P.mult[k];
    }
    double mv = d.mass[us];
    d.mass[u] = mv;
    double hv = d.hbuf[cur][us], rv = d.rbuf[cur][us];
    d.hbuf[0][u] = hv; d.hbuf[1][u] = hv;
    d.rbuf[0][u] = rv; d.rbuf[1][u] = rv;
    d.pmat[u] = d.pmat[us];
    d.gpar[gid] = from_real ? src : d.gpar[src];
    const int pm = (P.mult[0] < 0.0 ? 1 : 0) | (P.mult[1] < 0.0 ? 2 : 0) | (P.mult[2] < 0.0 ? 4 : 0);
    d.gsgn[gid] = (from_real ? 0 : d.gsgn[src]) ^ pm;
    // gather records of the ghost: start-of-step x, v, m, rho and h are
    // frozen for the whole step (mcm_ghost_setup.f90:330-341); the strain
    // pass does not shift ghosts (mcm_strain.f90:27-31), which is a shift
    // with a zero step, x - v 0 / 2, rounded as for a real particle
    const double sh = 0.0;
    const double s0 = d.x0[0][u] - 0.5 * d.v[0][u] * sh, s1 = d.x0[1][u] - 0.5 * d.v[1][u] * sh,
                 s2 = d.x0[2][u] - 0.5 * d.v[2][u] * sh;
    *rs_at(d, 0, u) = make_double2(s0, s1);
    *rs_at(d, 1, u) = make_double2(s2, hv);
    *rs_at(d, 2, u) = make_double2(mv / rv, d.v[0][u]);
    *rs_at(d, 3, u) = make_double2(d.v[1][u], d.v[2][u]);
    const double m0 = d.x0[0][u], m1 = d.x0[1][u], m2 = d.x0[2][u];
    *rm_at(d, 0, u) = make_double2(m0, m1);
    *rm_at(d, 1, u) = make_double2(m2, hv);
    *rm_at(d, 2, u) = make_double2(mv, mv / rv);
    const int okh = fast_h(hv);
    const int bad = ((fast_coord(s0) & fast_coord(s1) & fast_coord(s2) & okh) ? 0 : (int)RANGE_RS) |
                    ((fast_coord(m0) & fast_coord(m1) & fast_coord(m2) & okh) ? 0 : (int)RANGE_RM);
    if (bad) atomicOr(d.range, bad);
}

// Layer geometry of one plane (mcm_ghost_setup.f90:38-43 and siblings):
// cell r of the source layer S and the cell of the target layer T that its
// ghosts go to (1-based in the reference: symmetry copies layer 2 / lim-1,
// periodic the opposite end).
struct PlaneLayers {
    int S, T, dim, nlayer, L0, L1;
    __device__ __forceinline__ PlaneLayers(const PlaneDev& P, const GridDev& g) {
        dim = P.dim;
        const int lim = g.lim[dim];
        S = (P.side == 0 ? (P.code == 1 ? 2 : lim - 1) : (P.code == 1 ? lim - 1 : 2)) - 1;
        T = (P.side == 0 ? 1 : lim) - 1;
        L0 = g.lim[0];
        L1 = g.lim[1];
        nlayer = g.ncell / lim;
    }
    __device__ __forceinline__ void cells(int r, int& c, int& tc) const {
        int i, j, k;
        if (dim == 2) { i = r % L0; j = r / L0; k = S; }
        else if (dim == 1) { i = r % L0; k = r / L0; j = S; }
        else { j = r % L1; k = r / L1; i = S; }
        c = (k * L1 + j) * L0 + i;
        if (dim == 2) tc = (T * L1 + j) * L0 + i;
        else if (dim == 1) tc = (k * L1 + T) * L0 + i;
        else tc = (k * L1 + j) * L0 + T;
    }
};

// One plane of mcm_ghost_setup.f90:32-266, first half.  A single 1024-thread
// workgroup walks the source layer's cells in ascending cell id (the
// reference's loop order), counts each cell's members (its ghosts if it
// holds any, else its real particles) and scans the counts: gbase[r] is the
// id of the first ghost made from source cell r, in the reference's creation
// order.  Records each target cell's run of ghosts and the plane's ghost id
// range for k_ghost_fill.
__global__ __launch_bounds__(1024) void k_ghost_scan(DevState d_, int pi) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    Scalars* sc = d.sc;
    if (sc->abort) return;
    const PlaneLayers PL(d.planes[pi], sc->grid);
    __shared__ int s_ngp;
    __shared__ int s_over;
    if (threadIdx.x == 0) { s_ngp = sc->ngp; s_over = 0; }
    __syncthreads();
    const int lo = s_ngp;
    for (int r0 = 0; r0 < PL.nlayer; r0 += 1024) {
        const int r = r0 + threadIdx.x;
        int cnt = 0, c = -1, tc = -1;
        if (r < PL.nlayer) {
            PL.cells(r, c, tc);
            const int gsum = d.gcount[0][c] + d.gcount[1][c] + d.gcount[2][c];
            cnt = gsum > 0 ? gsum : d.cell_count[c];
        }
        int total;
        const int ex = block_excl_scan_1024(cnt, &total);
        const int base = s_ngp + ex;
        if (s_ngp + total > d.ngcap) {
            if (threadIdx.x == 0) s_over = 1;
        } else if (r < PL.nlayer) {
            d.gbase[r] = base;
            if (cnt > 0) {
                d.gstart[PL.dim][tc] = base;
                d.gcount[PL.dim][tc] = cnt;
                d.gany[tc] = 1;
                // a cell holding ghosts hides its real particles from the
                // search: no candidates of the skin filter this step
                if (d.skin > 0.0 && sc->skin_valid) {
                    double2* xh = reinterpret_cast<double2*>(d.xh);
                    const int s0 = d.cell_start[tc];
                    for (int q = 0; q < d.cell_count[tc]; q++) {
                        double* hq = &xh[2 * (long)d.skin_pos[d.sorted[s0 + q]] + 1].y;
                        *hq = -fabs(*hq);
                    }
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0 && !s_over) s_ngp += total;
        __syncthreads();
        if (s_over) break;
    }
    if (threadIdx.x == 0) {
        if (s_over) { sc->error_flags |= 2; sc->abort = 1; }
        sc->ngp = s_ngp;
        sc->gplane_lo = lo;
        sc->gplane_hi = s_ngp;
    }
}

// Second half: one thread per ghost of the plane.  The source cell of ghost
// g is the last layer cell r with gbase[r] <= g (cells without members share
// the next cell's base); within it the members are taken in the order the
// reference pushes them: a cell holding ghosts gives its ghosts of the z,
// then y, then x planes, each run from its last ghost down; a cell of real
// particles gives them in descending id (its sorted order).
__global__ __launch_bounds__(MCM_BLOCK) void k_ghost_fill(DevState d_, int pi, int cur) {
    Blk bk;
    const DevState& d = enter<false>(d_, bk);
    Scalars* sc = d.sc;
    if (sc->abort) return;
    const PlaneDev P = d.planes[pi];
    const PlaneLayers PL(P, sc->grid);
    const int lo = sc->gplane_lo, hi = sc->gplane_hi;
    for (int g = lo + bk.b * MCM_BLOCK + threadIdx.x; g < hi; g += bk.n * MCM_BLOCK) {
        int a = 0, b = PL.nlayer - 1;   // last r with gbase[r] <= g
        while (a < b) {
            const int mid = (a + b + 1) >> 1;
            if (d.gbase[mid] <= g) a = mid;
            else b = mid - 1;
        }
        int c, tc;
        PL.cells(a, c, tc);
        int m = g - d.gbase[a];
        const int gsum = d.gcount[0][c] + d.gcount[1][c] + d.gcount[2][c];
        if (gsum > 0) {
            for (int dd = 2; dd >= 0; dd--) {
                const int n = d.gcount[dd][c];
                if (m < n) {
                    make_ghost(d, cur, P, false, d.gstart[dd][c] + (n - 1 - m), g);
                    break;
                }
                m -= n;
            }
        } else {
            make_ghost(d, cur, P, true, d.sorted[d.cell_start[c] + m], g);
        }
    }
}

// ------------------------------------------------------------ list build
// What a skin build records of the particle i at sorted position t (cell c,
// x, h of this step): its position, the particle and material at t, x and h
// at the build, its block of the build, and this step's candidate record
// (h negated where a ghost-holding cell hides it; k_ghost_scan marks such
// particles in the steps between builds).  A block is clamped to the block
// grid: a clamped block is only nearer to the others, which keeps both skin
// conditions safe.
__device__ __forceinline__ void skin_record(const DevState& d, int t, int i, int c, double xi, double yi, double zi,
                                            double hi, const GridDev& g) {
    (void)g;
    const Scalars* sc = d.sc;
    d.skin_pos[i] = t;
    d.skin_id[t] = i;
    if (d.nummat > 1) d.skin_mat[t] = d.pmat[i];
    double2* br = reinterpret_cast<double2*>(d.skin_base) + 2 * (long)t;
    br[0] = make_double2(xi, yi);
    br[1] = make_double2(zi, hi);
    double2* xr = reinterpret_cast<double2*>(d.xh) + 2 * (long)t;
    xr[0] = make_double2(xi, yi);
    xr[1] = make_double2(zi, d.gany[c] ? -hi : hi);
    const double xv[3] = {xi, yi, zi};
    int kb[3];
    for (int a = 0; a < 3; a++) {
        const double q = (xv[a] - sc->skin_borig[a]) / sc->skin_bedge;
        const int nb = sc->skin_bdim[a];
        kb[a] = q >= 0.0 ? (q < (double)nb ? (int)q : nb - 1) : 0;
    }
    d.skin_blk[t] = (kb[2] * sc->skin_bdim[1] + kb[1]) * sc->skin_bdim[0] + kb[0];
}

// Ghost candidates of particle i (material mi, at x, h) in the cells
// [lo, hi] around its own (mcm_ll_neighbours.f90:181-228): the ghost runs
// of each ghost-holding cell, z then y then x planes, each from its last
// ghost down; accepted ghosts are appended to out, counted in cnt (gccnt:
// other-material ghosts of penalty contact, only counted), their distances
// folded into mind / mind_cp.
template <bool MULTI>
__device__ __forceinline__ void ghost_walk(const DevState& d, int cur, int i, int mi, double xi, double yi, double zi,
                                           double hi, int lo0, int hi0, int lo1, int hi1, int lo2, int hi2, int L0,
                                           int L1, bool penalty, int* out, int& cnt, int& gccnt, double& mind,
                                           double& mind_cp) {
    (void)i;
    const int ncap = d.ncap;
    for (int pass = 0; pass < (MULTI ? 2 : 1); pass++) {
        for (int n = lo2; n <= hi2; n++)
            for (int m = lo1; m <= hi1; m++)
                for (int l = lo0; l <= hi0; l++) {
                    int cc = (n * L1 + m) * L0 + l;
                    if (!d.gany[cc]) continue;
                    for (int dd = 2; dd >= 0; dd--) {
                        int gn = d.gcount[dd][cc];
                        if (gn == 0) continue;
                        int gs = d.gstart[dd][cc];
                        for (int q = gn - 1; q >= 0; q--) {
                            int u = d.np + gs + q;
                            double dist_sq = 0.0;
                            double dx = d.x0[0][u] - xi; dist_sq = dist_sq + dx * dx;
                            double dy = d.x0[1][u] - yi; dist_sq = dist_sq + dy * dy;
                            double dz = d.x0[2][u] - zi; dist_sq = dist_sq + dz * dz;
                            double h_avg = 2.0 * 0.5 * (d.hbuf[cur][u] + hi);
                            if (dist_sq < h_avg * h_avg) {
                                const bool same = !MULTI || d.pmat[u] == mi;
                                if (pass == 0 && same) {
                                    mind = fmin(mind, dist_sq);
                                    if (cnt < ncap) out[(long)cnt * 64] = u;
                                    cnt++;
                                } else if (MULTI && pass == 1 && !same) {
                                    mind_cp = fmin(mind_cp, dist_sq);
                                    if (penalty) {
                                        gccnt++;
                                    } else {
                                        if (cnt < ncap) out[(long)cnt * 64] = u;
                                        cnt++;
                                    }
                                }
                            }
                        }
                    }
                }
    }
}

// mcm_ll_neighbours.f90:96-299 for one real particle, one thread per
// particle (list layout: nbr_row in mcm_device.h).  MULTI: more than one
// material, so same-material and other-material passes.
//
// Real candidates are read from cell runs of the sorted order.  A pair is
// accepted only when its distance is below h_i + h_j <= h_i + hcap (hcap:
// largest h of the binned particles), so the real particles of a cell whose
// box lies at least that far away can never be accepted: such rows and
// x-run end cells are skipped (with a relative margin of 1e-9 that covers
// the rounding of the box faces).  Ghosts are not binned by position and
// are never skipped.  Skipping rejected candidates leaves the list, its
// order and mindist unchanged.
//
// The rows are visited in the same relative order by every lane (z, then y
// offsets -1, 0, 1), which is each particle's walk order.  For each row the
// lanes of a wavefront read overlapping runs of the sorted order (their
// particles sit in a few neighbouring cells), so the wavefront first copies
// the span of its lanes' runs into its own LDS slice, 32 bytes per
// candidate, and the lanes then walk their runs from there: every candidate
// record is fetched from the caches once per wavefront and row instead of
// once per lane.  A span longer than the slice (a wavefront that straddles a
// plane of cells) is staged and walked in pieces, in order.
//
// SKIN: the skin lists of a step that needs them (k_skin_check), with the
// same walk: every binned real particle j != i with
// |x_i - x_j| < (h_i + h_j + w)(1 + 1e-9), w the skin width, all materials
// in walk order, stored as sorted positions of this build; the real
// particles of ghost-holding cells are taken too (xh hides them in the
// steps where their cell holds ghosts).  Records each particle's sorted
// position, its x and h, its block and its candidate record.
#define LIST_STAGE 256   // candidates per wavefront slice (8 KB)
template <bool MULTI, bool SKIN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MULTI ? 3 : 4))) void k_list(DevState d_, int cur) {
    // threads take the binned particles in cell-sorted order, so that the
    // lanes of a wavefront share cells and read the same candidates
    // each slice has room for the three records a group of four reads past
    // a run's end (unused values)
    __shared__ double2 s_rec[4][2 * (LIST_STAGE + 4)];
    // the candidates' ids (and materials): SKIN lists hold sorted positions
    __shared__ int2 s_idm[MULTI && !SKIN ? 4 : 1][MULTI && !SKIN ? LIST_STAGE + 4 : 1];
    __shared__ int s_id[!MULTI && !SKIN ? 4 : 1][!MULTI && !SKIN ? LIST_STAGE + 4 : 1];
    Blk bk;
    const DevState& d = enter<true>(d_, bk);
    const int t = bk.b * blockDim.x + threadIdx.x;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    Scalars* sc = d.sc;
    if (sc->abort) return;
    if (SKIN && !sc->skin_rebuild) return;
    const GridDev g = sc->grid;
    const bool live = g.ncell > 0 && t < d.cell_start[g.ncell - 1] + d.cell_count[g.ncell - 1];
    const int L0 = g.lim[0], L1 = g.lim[1], L2 = g.lim[2];
    const int ncap = SKIN ? d.skin_cap : d.ncap;
    const int i = live ? d.sorted[t] : 0;
    const int c = live ? d.cell_of[i] : -1;
    const bool binned = live && c >= 0;
    int* out = SKIN ? skin_row(d, t) : nbr_row(d, t);
    // penalty contact: other-material partners go to the contact list, ghosts
    // of other materials are only counted (mcm_ll_neighbours.f90:162-174,
    // 209-223; merged into the lists only for kernel contact, :239-267)
    const bool penalty = !SKIN && MULTI && d.contacttype > 0;
    int cnt = 0, ccnt = 0, gccnt = 0;
    double mind = 1.0e20, mind_cp = 1.0e20;
    double xi = 0.0, yi = 0.0, zi = 0.0, hi = 0.0;
    int mi = 0, b0 = 0, b1 = 0, b2 = 0;
    const double w = SKIN ? sc->skin_len : 0.0;
    const double grow = SKIN ? 1.0 + 1.0e-9 : 1.0;
    double reach2 = 0.0;
    // distances from the particle to the lower and upper faces of its cell
    double flo[3] = {0.0, 0.0, 0.0}, fhi[3] = {0.0, 0.0, 0.0};
    if (binned) {
        xi = d.x0[0][i]; yi = d.x0[1][i]; zi = d.x0[2][i];
        hi = d.hbuf[cur][i];
        mi = d.pmat[i];
        b0 = c % L0; b1 = (c / L0) % L1; b2 = c / (L0 * L1);
        const double reach = (hi + __longlong_as_double((long long)sc->hcap_bits) + w) * grow * (1.0 + 1.0e-9);
        reach2 = reach * reach;
        const double xv[3] = {xi, yi, zi};
        const int bv[3] = {b0, b1, b2};
        for (int a = 0; a < 3; a++) {
            const double cw = 1.0 / g.gridsize[a];
            const double face = g.gridmin[a] + bv[a] * cw;
            flo[a] = fmax(xv[a] - face, 0.0);
            fhi[a] = fmax(face + cw - xv[a], 0.0);
        }
    }
    const double2* cd = reinterpret_cast<const double2*>(d.cand);
    const int2* ci = reinterpret_cast<const int2*>(d.cand_id);
    double2* rec = s_rec[wv];
    int2* idl = s_idm[MULTI && !SKIN ? wv : 0];
    int* idsl = s_id[!MULTI && !SKIN ? wv : 0];
    const int spos_i = t;
    // the walk's entries go to element oi of the list array (entry cnt of
    // this particle's row: rowbase + 64 cnt), clamped to the row's last slot:
    // a list that outgrows the capacity sets the error flag below and the
    // step is run again with rows long enough
    int* const lbase = SKIN ? d.skin_list : d.nbr;
    const unsigned rowbase = (unsigned)(t >> 6) * (unsigned)ncap * 64u + (unsigned)(t & 63);
    const unsigned lastoff = rowbase + (unsigned)(ncap - 1) * 64u;
    unsigned oi = rowbase;
    for (int pass = 0; pass < (MULTI && !SKIN ? 2 : 1); pass++) {
        auto append = [&](int v) {
            lbase[min(oi, lastoff)] = v;
            oi += 64u;
        };
        // candidate at sorted slot s with id and material idm (with one
        // material the self test uses the sorted slot)
        auto consider = [&](int s, double2 pa, double2 pb, int2 idm) {
            if (MULTI && !SKIN ? idm.x == i : s == spos_i) return;
            double dx = pa.x - xi;
            double dist_sq = dx * dx;   // 0.0 + dx*dx is dx*dx exactly
            double dy = pa.y - yi; dist_sq = dist_sq + dy * dy;
            double dz = pb.x - zi; dist_sq = dist_sq + dz * dz;
            if (SKIN) {
                const double r = (pb.y + hi + w) * grow;
                if (dist_sq < r * r) append(s);
                return;
            }
            double h_avg = 2.0 * 0.5 * (pb.y + hi);
            if (dist_sq < h_avg * h_avg) {
                const bool same = !MULTI || idm.y == mi;
                if (pass == 0 && same) {
                    mind = min_f64(mind, dist_sq);
                    append(idm.x);
                } else if (MULTI && pass == 1 && !same) {
                    mind_cp = min_f64(mind_cp, dist_sq);
                    if (penalty) {
                        if (ccnt < ncap) cont_list(d, i)[(long)ccnt * 64] = idm.x;
                        ccnt++;
                    } else {
                        append(idm.x);
                    }
                }
            }
        };
        // candidates [st, en) of the slice holding [base, base +
        // LIST_STAGE), four at a time: the loads of a group are issued
        // together (past the run's end they read slots whose values go
        // unused), the tests run in walk order
        auto walk = [&](int st, int en, int base) {
            const double2* r = rec + 2 * (st - base);
            const int2* pm = idl + (st - base);
            const int* pi = idsl + (st - base);
            auto group = [&](int s, bool tail) {
                int2 idm[4];
                double2 pa[4], pb[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (MULTI && !SKIN) idm[u] = pm[u];
                    if (!MULTI && !SKIN) idm[u].x = pi[u];
                    pa[u] = r[2 * u];
                    pb[u] = r[2 * u + 1];
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (!tail || s + u < en) consider(s + u, pa[u], pb[u], idm[u]);
            };
            int s = st;
            for (; s + 4 <= en; s += 4, r += 8, pm += 4, pi += 4) group(s, false);
            if (s < en) group(s, true);
        };
#pragma unroll 1
        for (int dz = -1; dz <= 1; dz++) {
#pragma unroll 1
            for (int dy = -1; dy <= 1; dy++) {
                const int n = b2 + dz, m = b1 + dy;
                const double fz = dz < 0 ? flo[2] : (dz > 0 ? fhi[2] : 0.0);
                const double fy = dy < 0 ? flo[1] : (dy > 0 ? fhi[1] : 0.0);
                const double dyz = fz * fz + fy * fy;
                // this lane's run: the row's cells b0 - 1 .. b0 + 1, less the
                // end cells and the row out of reach
                const bool on = binned && n >= 0 && n < L2 && m >= 0 && m < L1 && dyz < reach2;
                int l0 = b0, l1 = b0, st = 0, en = 0;
                const int c0 = (n * L1 + m) * L0;
                bool ghosty = false;
                if (on) {
                    if (b0 > 0 && dyz + flo[0] * flo[0] < reach2) l0 = b0 - 1;
                    if (b0 + 1 < L0 && dyz + fhi[0] * fhi[0] < reach2) l1 = b0 + 1;
                    st = d.cell_start[c0 + l0];
                    en = d.cell_start[c0 + l1] + d.cell_count[c0 + l1];
                    // a row holding a ghost-headed cell is walked cell by cell
                    // so that that cell's real particles stay hidden
                    if (!SKIN)
                        for (int l = l0; l <= l1; l++) ghosty |= d.gany[c0 + l] != 0;
                }
                // the wavefront's span of runs
                int lo = st < en ? st : INT_MAX, hi_s = st < en ? en : INT_MIN;
                for (int off = 32; off > 0; off >>= 1) {
                    lo = min(lo, __shfl_xor(lo, off, 64));
                    hi_s = max(hi_s, __shfl_xor(hi_s, off, 64));
                }
                // the span in slices, each walked by the lanes over their
                // runs' share of it, in order
                for (int cb = lo; cb < hi_s; cb += LIST_STAGE) {
                    const int ce = min(cb + LIST_STAGE, hi_s);
                    __builtin_amdgcn_wave_barrier();       // the slice's last readers are done
                    for (int k = lane; k < ce - cb; k += 64) {
                        rec[2 * k] = cd[2 * (cb + k)];
                        rec[2 * k + 1] = cd[2 * (cb + k) + 1];
                        if (MULTI && !SKIN) idl[k] = ci[cb + k];
                        if (!MULTI && !SKIN) idsl[k] = ci[cb + k].x;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (!on) continue;
                    if (!ghosty) {
                        walk(max(st, cb), min(en, ce), cb);
                    } else {
                        for (int l = l0; l <= l1; l++) {
                            const int cc = c0 + l;
                            if (d.gany[cc]) continue;
                            const int cs = d.cell_start[cc];
                            walk(max(cs, cb), min(cs + d.cell_count[cc], ce), cb);
                        }
                    }
                }
            }
        }
    }
    cnt = (int)((oi - rowbase) >> 6);
    if (live) {
        int nreal;
        if (binned) {
            if (SKIN) {
                skin_record(d, t, i, c, xi, yi, zi, hi, g);
            } else {
                if (d.disctype == 1) {
                    if (cnt < ncap) out[(long)cnt * 64] = i;
                    cnt++;
                }
            }
            nreal = cnt;
            if (!SKIN && d.boundary) {
                const int lo0 = max(b0 - 1, 0), hi0 = min(b0 + 1, L0 - 1);
                const int lo1 = max(b1 - 1, 0), hi1 = min(b1 + 1, L1 - 1);
                const int lo2 = max(b2 - 1, 0), hi2 = min(b2 + 1, L2 - 1);
                ghost_walk<MULTI>(d, cur, i, mi, xi, yi, zi, hi, lo0, hi0, lo1, hi1, lo2, hi2, L0, L1, penalty, out,
                                  cnt, gccnt, mind, mind_cp);
            }
            mind = sqrt(penalty ? mind : fmin(mind, mind_cp));
            // a particle with real and ghost contact partners: the reference
            // reads the ghost partners from the real contact list past its end
            // (mcm_get_j_cont_info.f90:44-48), so its repulsion is undefined
            if (penalty && ccnt > 0 && gccnt > 0) {
                atomicOr(&sc->error_flags, 32);
                atomicExch(&sc->abort, 1);
            }
        } else {
            if (!SKIN && d.disctype == 1) {
                if (cnt < ncap) out[0] = i;
                cnt++;
            }
            nreal = cnt;
        }
        if (SKIN) {
            d.skin_n[i] = cnt;
        } else {
            d.nreal[i] = nreal;
            d.ntot[i] = cnt;
            d.mindist[i] = mind;
            if (penalty) d.ncont[i] = ccnt;
            cnt = max(cnt, ccnt);   // both lists share the capacity ncap
        }
    }
    // longest list of the wavefront, one atomic per wavefront
    for (int off = 32; off > 0; off >>= 1) cnt = max(cnt, __shfl_xor(cnt, off, 64));
    if ((threadIdx.x & 63) == 0 && cnt > 0) {
        atomicMax(SKIN ? &sc->skin_max : &sc->max_count, cnt);
        if (cnt > (SKIN ? d.skin_cap : d.ncap)) {
            atomicOr(&sc->error_flags, SKIN ? 64 : 1);
            atomicExch(&sc->abort, 1);
        }
    }
    if (SKIN && t == 0) sc->skin_nbuilt = live ? d.cell_start[g.ncell - 1] + d.cell_count[g.ncell - 1] : 0;
}

// ----------------------------------------------------- skin list filter
// The neighbour lists of one step from the skin lists (see the skin notes
// above): each binned particle, in this step's cell-sorted thread order,
// keeps the entries of its skin list that pass the reference's test
// (mcm_ll_neighbours.f90:140-148, the expression of k_list) with x and h of
// this step, same material first, then other materials (kernel contact),
// in skin-list order; then itself (corrected SPH) and its ghosts as k_list
// finds them.  The sets equal k_list's; the order of the real partners is
// the skin build's.  Runs with penalty contact take no skin.
template <bool MULTI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MULTI ? 3 : 5))) void k_skin_filter(DevState d_,
                                                                                                        int cur) {
    Blk bk;
    const DevState& d = enter<true>(d_, bk);
    const int t = bk.b * blockDim.x + threadIdx.x;
    Scalars* sc = d.sc;
    if (sc->abort) return;
    const GridDev g = sc->grid;
    const bool live = g.ncell > 0 && t < d.cell_start[g.ncell - 1] + d.cell_count[g.ncell - 1];
    int cnt = 0;
    if (live) {
        const int L0 = g.lim[0], L1 = g.lim[1];
        const int ncap = d.ncap;
        int* out = nbr_row(d, t);
        const double2* cd = reinterpret_cast<const double2*>(d.cand);
        const int2 own = reinterpret_cast<const int2*>(d.cand_id)[t];
        const int i = own.x, mi = own.y;
        const int c = d.cell_of[i];
        const double2 pa0 = cd[2 * t], pb0 = cd[2 * t + 1];
        const double xi = pa0.x, yi = pa0.y, zi = pb0.x, hi = pb0.y;
        const int* sl = skin_row(d, d.skin_pos[i]);
        const int sn = d.skin_n[i];
        const double2* xh = reinterpret_cast<const double2*>(d.xh);
        double mind = 1.0e20, mind_cp = 1.0e20;
        for (int pass = 0; pass < (MULTI ? 2 : 1); pass++) {
            for (int k = 0; k < sn; k += 4) {
                int j[4];
                double2 pa[4], pb[4];
#pragma unroll
                for (int u = 0; u < 4; u++) j[u] = sl[(long)min(k + u, sn - 1) * 64];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    pa[u] = xh[2 * (long)j[u]];
                    pb[u] = xh[2 * (long)j[u] + 1];
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    if (k + u >= sn) continue;
                    double dx = pa[u].x - xi;
                    double dist_sq = dx * dx;   // 0.0 + dx*dx is dx*dx exactly
                    double dy = pa[u].y - yi; dist_sq = dist_sq + dy * dy;
                    double dz = pb[u].x - zi; dist_sq = dist_sq + dz * dz;
                    double h_avg = 2.0 * 0.5 * (pb[u].y + hi);
                    if (pb[u].y > 0.0 && dist_sq < h_avg * h_avg) {
                        const bool same = !MULTI || d.skin_mat[j[u]] == mi;
                        if (pass == 0 && same) {
                            mind = fmin(mind, dist_sq);
                            if (cnt < ncap) out[(long)cnt * 64] = d.skin_id[j[u]];
                            cnt++;
                        } else if (MULTI && pass == 1 && !same) {
                            mind_cp = fmin(mind_cp, dist_sq);
                            if (cnt < ncap) out[(long)cnt * 64] = d.skin_id[j[u]];
                            cnt++;
                        }
                    }
                }
            }
        }
        if (d.disctype == 1) {
            if (cnt < ncap) out[(long)cnt * 64] = i;
            cnt++;
        }
        const int nreal = cnt;
        if (d.boundary) {
            const int b0 = c % L0, b1 = (c / L0) % L1, b2 = c / (L0 * L1);
            const int lo0 = max(b0 - 1, 0), hi0 = min(b0 + 1, L0 - 1);
            const int lo1 = max(b1 - 1, 0), hi1 = min(b1 + 1, L1 - 1);
            const int lo2 = max(b2 - 1, 0), hi2 = min(b2 + 1, g.lim[2] - 1);
            int gccnt = 0;
            ghost_walk<MULTI>(d, cur, i, mi, xi, yi, zi, hi, lo0, hi0, lo1, hi1, lo2, hi2, L0, L1, false, out, cnt,
                              gccnt, mind, mind_cp);
        }
        d.nreal[i] = nreal;
        d.ntot[i] = cnt;
        d.mindist[i] = sqrt(fmin(mind, mind_cp));
    }
    for (int off = 32; off > 0; off >>= 1) cnt = max(cnt, __shfl_xor(cnt, off, 64));
    if ((threadIdx.x & 63) == 0 && cnt > 0) {
        atomicMax(&sc->max_count, cnt);
        if (cnt > d.ncap) {
            atomicOr(&sc->error_flags, 1);
            atomicExch(&sc->abort, 1);
        }
    }
}
