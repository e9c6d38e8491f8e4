/* mcm_oracle.h -- CPU restatement of the MCM SPH time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is checked
 * against (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg).
 * Nothing in the product links it.
 *
 * It restates, in plain C over structure-of-arrays state, the reference's
 * initialisation chain and every stage of one time step, keeping each
 * floating-point expression in the reference's evaluation order so that it
 * reproduces the flang build of the reference bit for bit on the same deck
 * (pinned by tests/test_oracle_golden.py against SHA-256 digests of the
 * reference's own state dumps, and by tests/test_oracle_closed_form.py
 * against closed forms and brute force).  Each function cites the
 * reference routine it follows (MCM/<file>.f90:<lines>).
 *
 * Every branch restated here is exercised by a reference digest, except
 * three modes of the cap model f3dm25 (material model 4): parity unpinned
 * for modes 6 and 7 (0 < capa < 1e-7 fe) and 8 (capa < 0), which no deck
 * reaches (mcm_f3dm25.f90:229-261; capa stays >= 0 from its zero start).  Penalty
 * contact with ghost contacts is not restated: the reference reads its ghost contact partners
 * from the real contact list past its end (mcm_get_j_cont_info.f90:44-48),
 * so its result is undefined there; rep_cont stops with ORC_ERR_UNDEFINED.
 *
 * Conventions: 0-based particle and ghost indices; 3x3 tensors stored
 * column-major, element (r,c) at [9*i + r + 3*c]; vectors at [3*i + d].
 */
#ifndef MCM_ORACLE_H
#define MCM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int model, eos, visc_type;
    double rho, g, av_l, av_q, mass, h, rho_min, rho_max;
    double strinput[48];
    double eosinput[24];
    double k_cont, n_cont;      /* penalty contact K and n (mcm_contact_in.f90) */
} orc_material;

typedef struct {
    /* control */
    int np, nummat, disctype, h_opt, tcrit_opt, contacttype, krtype;
    int massopt, init_h_opt, init_rhoe, boundary;
    int veloc_opt;              /* 1: XSPH velocity smoothing */
    int ncontmats, cont_opt;    /* penalty contact cards */
    int drelax, baseaccel;      /* dynamic relaxation; base acceleration */
    double drelax_scale, base_a[3];
    int boundary_code[3][2];   /* [dim][0=min,1=max]: 0 none, 1 symmetry, 2 periodic */
    int boundary_type[3];
    double boundary_x[3][2];
    double tssf, itss, endtime, pi;
    /* time */
    double dt, dtold, ptime, critts, hmax;
    int timestep;
    /* grid */
    double coord_maxmin[3][2]; /* [dim][0=min,1=max] */
    double gridmin[3], gridsize[3];
    int gridlim[3];
    int *llgrid;                /* gridlim product; >0 real id+1, <0 -(ghost id+1), 0 empty */
    long llgrid_cap;
    double plane_mult[6][3], plane_add[6][3];  /* xmin xmax ymin ymax zmin zmax */
    orc_material *mat;
    /* particles */
    int *pmat, *dispbc, *llpointer;
    double *x, *xzero, *v, *a;                 /* 3 per particle */
    double *mass, *h, *hold, *h0, *rho, *rho0, *rhoold, *p, *e, *etry, *einc, *c, *vabs;
    double *tracerod, *pcut, *p_cut, *efps, *fail, *mindist, *critts_p;
    double *sigma, *s, *alfa, *q, *qold, *rod, *spin; /* 9 per particle */
    double *rep;                /* 3 per particle: contact acceleration of mcm_rep_cont */
    double *capa;               /* cap position of the cap model (material model 4) */
    int *ptype;                 /* mode taken by the cap model in its last update */
    int cap_imode;              /* the cap model's saved mode variable (a static local there) */
    /* neighbour lists (0-based ids), capacity ncap / gcap per particle */
    int ncap, gcap;
    int *nnbr, *g_nnbr, *ncont, *g_ncont;
    int *nbr, *gnbr, *cont, *gcont;
    /* ghosts */
    int ngp, ngp_cap;
    int *g_par, *g_mat, *g_ll;
    double *g_x, *g_xzero, *g_v, *g_mass, *g_h, *g_hold, *g_rho, *g_sigma, *g_q;
    /* energies */
    double *mat_ke, *mat_ie, *mat_mom;
    double kinetice, internale, totale;
    /* f3dm10 keeps its plastic-strain increment between calls when a particle
       has failed (the reference reads it uninitialised, mcm_f3dm10.f90:82-89) */
    double f3dm10_defps;
    int error;                 /* first error code seen */
} orc_state;

/* error codes */
#define ORC_OK 0
#define ORC_ERR_CAPACITY -2
#define ORC_ERR_MODEL -3
#define ORC_ERR_IO -4
#define ORC_ERR_UNDEFINED -5   /* penalty contact with ghost contacts: the reference reads an undefined list entry */

orc_state *orc_load_deck(const char *path, int ncap, int gcap);
void orc_free(orc_state *st);
int orc_initial(orc_state *st);
int orc_neighbours(orc_state *st);
int orc_strain(orc_state *st);
int orc_update_h(orc_state *st);
int orc_constitutive(orc_state *st);
int orc_dt_update(orc_state *st);
int orc_rep_cont(orc_state *st);
int orc_momentum(orc_state *st);
int orc_boundcond(orc_state *st);
int orc_calctotale(orc_state *st);
int orc_update_velocity(orc_state *st);
int orc_move(orc_state *st);
int orc_step(orc_state *st);
int orc_run(orc_state *st, int nsteps);
int orc_dump(orc_state *st, const char *path, int step, int stage);

/* positions (3 per particle) of the current state, for finite-difference
   checks of the GPU adjoint */
void orc_get_x(const orc_state *st, double *out);
void orc_get_rows(const orc_state *st, double *rows);
int orc_get_lists(const orc_state *st, int32_t *counts, int32_t *gcounts, int32_t *ids, int64_t cap);
void orc_get_scalars(const orc_state *st, double *out);
/* cap position and last mode of the cap model, per particle */
void orc_get_capa(const orc_state *st, double *capa, int32_t *ptype);

/* closed-form pieces exposed for unit tests */
double orc_kernel_w(const double *xi, const double *xj, double havg, double pi);
void orc_kernel_gradw(const double *xi, const double *xj, double havg, double pi, double *dwdx);
void orc_inverse3(const double *m, double *inv);

#ifdef __cplusplus
}
#endif
#endif
