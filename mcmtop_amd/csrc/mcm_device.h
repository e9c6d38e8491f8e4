// mcm_device.h -- device-resident state of one MCM run and the scalar block
// shared by every kernel of a time step.
#pragma once

#include "mcm_physics.h"

#define MCM_MAX_MAT 8
#define MCM_BLOCK 256

// Bits of DevState::range, set by the kernels that write the pair records
// when a coordinate or h lies outside the ranges of the pair loops' fast
// path (fast_coord, fast_h), cleared when a neighbour search starts:
// RANGE_RS for the strain records rs, RANGE_RM for the momentum records rm.
enum { RANGE_RS = 1, RANGE_RM = 2 };

// butterfly reductions over the 64 lanes of a wavefront (all lanes active)
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

// gfx950 dispatches the workgroups of a launch to its 8 XCDs round-robin
// (workgroup b runs on XCD b % 8), and each XCD has its own L2.  Gather
// kernels renumber their workgroups so that every XCD walks one contiguous
// range of particles: particles that share neighbours then share the cached
// neighbour records in one L2 instead of fetching them into all eight.
__device__ __forceinline__ int xcd_block() {
    const int b = blockIdx.x, nb = gridDim.x;
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Grid of the cell search (mcm_setuplist.f90:30-91); cell indices 1-based
// as in the reference, linear id = ((k-1)*lim1 + (j-1))*lim0 + (i-1).
struct GridDev {
    double gridmin[3], gridsize[3];
    int lim[3];
    int ncell;
};

// Scalars that the reference keeps in module variables, kept on the device
// so that a step needs no host round trip.
struct Scalars {
    double dt, dtold, ptime, critts, hmax;
    double coord_min[3], coord_max[3];
    double kinetice, internale, totale;
    double mat_ke[MCM_MAX_MAT], mat_ie[MCM_MAX_MAT], mat_mom[MCM_MAX_MAT][3];
    GridDev grid;
    int ngp;            // ghosts generated this step
    int nout;           // real particles outside the search grid
    int gplane_lo, gplane_hi; // ghost ids made by the plane being generated
    unsigned long long hcap_bits;  // bit pattern of the largest h among the binned particles
    int max_count;      // largest neighbour count seen by the last search
    int error_flags;    // MCM error bits, see mcm_gpu.h
    int abort;          // set when a capacity was exceeded: later kernels of the step do nothing
    int timestep;
    // neighbour skin (DevState::skin > 0): the displacement of all real
    // particles since the skin lists were built, as ordered bit patterns
    // (ordbits) of the per-axis minimum and maximum of x - x_build and the
    // maximum of |h - h_build|; whether two adjacent blocks moved too far
    // apart (skin_near); the skin width and the largest h of the build; the
    // block grid of the build (origin, edge, blocks per axis); whether the
    // lists are valid, are built in this step, how often they were built,
    // the longest skin list and the particles of the build
    unsigned long long skin_dmin[3], skin_dmax[3], skin_dh;
    int skin_near;
    double skin_len, skin_hcap;
    double skin_borig[3], skin_bedge;
    int skin_bdim[3];
    int skin_valid, skin_rebuild, skin_builds, skin_max, skin_nbuilt;
};

// Doubles as unsigned integers with the same order (for atomicMin / atomicMax
// on signed values), and back.
__host__ __device__ __forceinline__ unsigned long long ordbits(double x) {
    unsigned long long b;
    __builtin_memcpy(&b, &x, 8);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ double ordval(unsigned long long b) {
    b = (b >> 63) ? (b & 0x7fffffffffffffffull) : ~b;
    double x;
    __builtin_memcpy(&x, &b, 8);
    return x;
}

// Boundary planes in the order xmin xmax ymin ymax zmin zmax
struct PlaneDev {
    double mult[3], add[3];
    int dim, side, code, pad_;
};

// One workgroup of a batched launch: the instance it works for, its block
// number within that instance and the instance's block count in the launch.
struct BlockEntry {
    int inst, block, nblocks, pad_;
};

// Device pointers of one run.  Arrays marked [U] span the unified index
// space of np real particles followed by the ghost slots.
struct DevState {
    int np, npad, ngcap, nummat, disctype, h_opt, tcrit_opt, boundary, ncap;
    int nplanes, cellcap, need_s, has_model9;
    int nblk;                  // workgroups of MCM_BLOCK threads over np (per-particle launches)
    int contacttype;           // > 0: penalty contact (separate contact lists, mcm_rep_cont)
    int veloc_opt;             // 1: XSPH velocity smoothing
    int drelax, baseaccel;     // dynamic relaxation; base acceleration
    double drelax_scale, base_a[3];
    int bcode[3][2];
    int btype[3];
    double bx[3][2];
    double tssf, pi;
    PlaneDev planes[6];
    MatParams* mat;            // [nummat]
    Scalars* sc;
    // particle state
    double *x0[3], *x1[3];     // [U] x^n (strain input) and x after the strain pass
    double *v[3];              // [U]
    double* mass;              // [U]
    double *hbuf[2], *rbuf[2]; // [U] h and rho, double buffered (cur = step parity)
    int* pmat;                 // [U] 0-based material
    int* dispbc;
    int* gpar;                 // [ngcap] parent of each ghost (0-based real id)
    int* gsgn;                 // [ngcap] bit k set: the ghost's x_k and v_k are its parent's mirrored
                               // an odd number of times (product of the plane multipliers is -1)
    double *hold, *h0, *rho0, *p, *e, *c, *vabs, *pcut, *p_cut, *efps, *fail, *q, *mindist;
    double *sig[6], *sd[6], *alfa[6];
    double* rs;                // [U][8] strain gather record {x, v, m/rho, h}: x the strain pass's
                               // position (x - v dt / 2 for real particles), the rest start of step
    double* rm;                // [U][12] momentum gather record {x, h, m, m/rho, (sigma - q)/rho^2}
    // neighbour search
    int *cell_of, *slot_of, *cell_count, *cell_start, *sorted;
    int* unsorted;             // [np] cell runs in histogram slot order (k_scatter), sorted by k_sort_cells
    double* cand;              // [np][4] cell-sorted {x, y, z, h} of the candidates
    int* cand_id;              // [np][2] cell-sorted {particle id, material}
    int* spos;                 // [np] sorted slot of each binned particle
    int *gstart[3], *gcount[3];
    int* gbase;                // [cellcap] first ghost id of each source-layer cell of a plane
    int* gany;                 // [cellcap] 1 where a cell holds ghosts
    int* nbr;                  // [npad/64][ncap][64]: entry k of particle i at ((i/64)*ncap + k)*64 + i%64
    int *nreal, *ntot;
    // neighbour skin (skin > 0, skin = width as a fraction of the largest h):
    // real partners within h_i + h_j + width at the last build, blocked like
    // nbr by the particle's sorted position at that build
    double skin;
    int skin_cap, skin_blkcap;
    int* skin_list;            // [npad/64][skin_cap][64]
    int* skin_n;               // [np] entries of each particle's skin list
    int* skin_pos;             // [np] sorted position of each particle at the build
    int* skin_id;              // [np] particle at each sorted position of the build
    int* skin_blk;             // [np] block of the particle at each sorted position of the build
    int* skin_mat;             // [np] material of the particle at each sorted position of the build
    double* skin_base;         // [np][4] x, y, z, h at the build, by sorted position of the build
    unsigned long long* skin_bstat[2]; // [skin_blkcap][8] per block: ordbits of min x - x_build (3), max (3),
                               // max |h - h_build|; one buffer filled per step, the other cleared
    double* skin_red;          // [skin_nred][8] per-workgroup extents of k_skin_blocks
    int skin_nred;             // workgroups of k_skin_blocks
    double* xh;                // [np][4] x, y, z, h this step by sorted position of the skin build
                               // (h negated: no candidate this step)
    int* cont;                 // penalty contact lists, blocked like nbr (contacttype > 0)
    int* ncont;                // [np] real contact partners
    double* rep[3];            // [np] contact acceleration of mcm_rep_cont
    double* sv[3];             // [np] XSPH smoothing of the velocity
    int* scan_tmp;
    // diagnostics / stage outputs
    double *rod[6], *spin[3], *tracerod, *a[3], *einc, *etry, *rhoold, *critts_p;
    int* yield;                // active-set mask of the plastic models
    double* red;               // per-block partial reductions
    int* range;                // RANGE_* bits: records outside the pair loops' fast-path ranges
    double* defps_prev;        // f3dm10 increment carried for failed particles
    // the cap model (material model 4) writes only the upper triangle of the
    // stress (mcm_f3dm25.f90:801-806): runs with a cap-model material (asym)
    // keep the lower triangle too, and the momentum passes read all nine
    // components
    int asym;
    double* capa;              // [np] cap parameter
    int* ptype;                // [np] mode of the last cap-model update
    double* sigl[3];           // [np] sigma 21, 31, 32 (the upper triangle's mirror for other models)
    double* rml;               // [U][4] (sigma 21, 31, 32) / rho^2 of the momentum record, and a pad
    // batched launch (null for a launch over one run): workgroup -> instance
    // block, and the states of all instances
    const BlockEntry* bmap;
    const DevState* states;
};

// The workgroup of this launch as one run sees it: block b of n.  A launch
// over one run uses the grid itself (XCD: renumbered by xcd_block) and the
// state passed as the kernel argument; a batched launch looks its workgroup
// up in the block map and reads the state of the instance it works for.
// Both states are read through the constant address space, so their fields
// stay scalar loads.  Every kernel that a batched step launches calls this
// first.
struct Blk {
    int b, n;
};
typedef const __attribute__((address_space(4))) DevState ConstDevState;
template <bool XCD>
__device__ __forceinline__ const DevState& enter(const DevState& d0, Blk& bk) {
    (void)d0;   // the kernel's first argument, read below through the kernarg segment
    const int g = XCD ? xcd_block() : (int)blockIdx.x;
    ConstDevState* p = (ConstDevState*)__builtin_amdgcn_kernarg_segment_ptr();
    if (p->bmap == nullptr) {
        bk = Blk{g, (int)gridDim.x};
    } else {
        const BlockEntry e = p->bmap[g];
        bk = Blk{e.block, e.nblocks};
        p = (ConstDevState*)p->states + e.inst;
    }
    return *(const DevState*)p;
}

// Neighbour lists are stored by position in the cell-sorted order of the
// search (sorted[t] is the particle at position t and spos[i] the position
// of particle i; particles outside the grid take the last positions), in
// blocks of 64 positions: for each block, slot k of all 64 is one
// contiguous 256-byte row, so a wavefront of threads reading slot k of its
// positions reads one row.  The pair kernels take the particles in this
// order: the lanes of a wavefront then lie in a few adjacent cells and
// gather neighbours from fewer cache lines than particles taken in id
// order.  Entry k of the particle at position t is nbr_row(d, t)[k * 64].
__device__ __forceinline__ int* nbr_row(const DevState& d, int t) {
    return d.nbr + (long)(t >> 6) * d.ncap * 64 + (t & 63);
}
// entry k of the skin list of the particle at sorted position t of the build
__device__ __forceinline__ int* skin_row(const DevState& d, int t) {
    return d.skin_list + (long)(t >> 6) * d.skin_cap * 64 + (t & 63);
}
// Pair gather records, one aligned record per entry of the unified index
// space, in 16-byte slots:
//   rs[u] (64 bytes): (x, y), (z, h), (m/rho, vx), (vy, vz)
//   rm[u] (96 bytes): (x, y), (z, h), (m, m/rho), (A11, A22), (A33, A12), (A13, A23)
// The pair loops' gathers are bound by the vector L1 path (the lanes of a
// wavefront read 64 different neighbours per step).  Measured on the
// 2.1 M-particle bar: whole records beat one array per slot (more lines per
// neighbour, slower once the bar deforms), and the 96-byte momentum record
// beats one padded to a 128-byte line (fewer L2 requests, but no two
// lanes then share a line).
enum { RS_SLOTS = 4, RM_SLOTS = 6 };
__device__ __forceinline__ double2* rs_at(const DevState& d, int p, long u) {
    return reinterpret_cast<double2*>(d.rs) + RS_SLOTS * u + p;
}
__device__ __forceinline__ double2* rm_at(const DevState& d, int p, long u) {
    return reinterpret_cast<double2*>(d.rm) + RM_SLOTS * u + p;
}
// the first double of a slot
__device__ __forceinline__ double* lo_of(double2* p) { return reinterpret_cast<double*>(p); }

__device__ __forceinline__ int* cont_list(const DevState& d, int i) {
    return d.cont + (long)(i >> 6) * d.ncap * 64 + (i & 63);
}
