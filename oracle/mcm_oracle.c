/* mcm_oracle.c -- CPU restatement of the MCM SPH time step.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  See mcm_oracle.h.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (oracle/Makefile).  FP
 * contraction must stay off: every expression below is written in the
 * reference's left-to-right evaluation order so that results agree bit for
 * bit with the flang build of the reference on the same deck.
 *
 * Single-precision literals of the reference are spelled as (double)(float)
 * where their value is not exact in binary (1.0e-3, 1.e-30, 1.e-12, 1.e30).
 */
#include "mcm_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define T9(a, i, r, c) ((a)[9 * (long)(i) + (r) + 3 * (c)])
#define V3(a, i, d) ((a)[3 * (long)(i) + (d)])
#define CM(st, m, a, b) ((st)->mat[m].strinput[((a) - 1) + 5 * ((b) - 1)])

/* The pair loops (list build, strain, acceleration, contact) run their
   particles on OpenMP threads: each particle's sums stay in its own thread,
   in list order, so the results do not depend on the thread count. */
static void set_err(orc_state *st, int code) {
#pragma omp critical(orc_error)
    if (st->error == 0) st->error = code;
}

/* Fortran SIGN(1.0, x) */
static double fsign1(double x) { return signbit(x) ? -1.0 : 1.0; }
static double fdim_(double a, double b) { return a - b > 0.0 ? a - b : 0.0; }
static double fmax_(double a, double b) { return a > b ? a : b; }
static double fmin_(double a, double b) { return a < b ? a : b; }

/* ------------------------------------------------------------------ memory */
static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

static void alloc_particles(orc_state *st) {
    long n = st->np;
    st->pmat = xcalloc(n, sizeof(int));
    st->dispbc = xcalloc(n, sizeof(int));
    st->llpointer = xcalloc(n, sizeof(int));
    st->x = xcalloc(3 * n, 8); st->xzero = xcalloc(3 * n, 8);
    st->v = xcalloc(3 * n, 8); st->a = xcalloc(3 * n, 8);
    double **s1[] = {&st->mass, &st->h, &st->hold, &st->h0, &st->rho, &st->rho0, &st->rhoold, &st->p,
                     &st->e, &st->etry, &st->einc, &st->c, &st->vabs, &st->tracerod, &st->pcut,
                     &st->p_cut, &st->efps, &st->fail, &st->mindist, &st->critts_p};
    for (size_t k = 0; k < sizeof(s1) / sizeof(s1[0]); k++) *s1[k] = xcalloc(n, 8);
    st->rep = xcalloc(3 * n, 8);
    st->capa = xcalloc(n, 8);
    st->ptype = xcalloc(n, sizeof(int));
    double **s9[] = {&st->sigma, &st->s, &st->alfa, &st->q, &st->qold, &st->rod, &st->spin};
    for (size_t k = 0; k < sizeof(s9) / sizeof(s9[0]); k++) *s9[k] = xcalloc(9 * n, 8);
    st->nnbr = xcalloc(n, sizeof(int)); st->g_nnbr = xcalloc(n, sizeof(int));
    st->ncont = xcalloc(n, sizeof(int)); st->g_ncont = xcalloc(n, sizeof(int));
    st->nbr = xcalloc((size_t)n * st->ncap, sizeof(int));
    st->cont = xcalloc((size_t)n * st->ncap, sizeof(int));
    st->gnbr = xcalloc((size_t)n * st->gcap, sizeof(int));
    st->gcont = xcalloc((size_t)n * st->gcap, sizeof(int));
    st->ngp_cap = 4 * st->np + 1024;
    long g = st->ngp_cap;
    st->g_par = xcalloc(g, sizeof(int)); st->g_mat = xcalloc(g, sizeof(int)); st->g_ll = xcalloc(g, sizeof(int));
    st->g_x = xcalloc(3 * g, 8); st->g_xzero = xcalloc(3 * g, 8); st->g_v = xcalloc(3 * g, 8);
    st->g_mass = xcalloc(g, 8); st->g_h = xcalloc(g, 8); st->g_hold = xcalloc(g, 8); st->g_rho = xcalloc(g, 8);
    st->g_sigma = xcalloc(9 * g, 8); st->g_q = xcalloc(9 * g, 8);
}

void orc_free(orc_state *st) {
    if (!st) return;
    void *ptrs[] = {st->pmat, st->dispbc, st->llpointer, st->x, st->xzero, st->v, st->a, st->mass, st->h,
                    st->hold, st->h0, st->rho, st->rho0, st->rhoold, st->p, st->e, st->etry, st->einc, st->c,
                    st->vabs, st->tracerod, st->pcut, st->p_cut, st->efps, st->fail, st->mindist,
                    st->critts_p, st->sigma, st->s, st->alfa, st->q, st->qold, st->rod, st->spin, st->nnbr,
                    st->g_nnbr, st->ncont, st->g_ncont, st->nbr, st->cont, st->gnbr, st->gcont, st->g_par,
                    st->g_mat, st->g_ll, st->g_x, st->g_xzero, st->g_v, st->g_mass, st->g_h, st->g_hold,
                    st->g_rho, st->g_sigma, st->g_q, st->mat, st->mat_ke, st->mat_ie, st->mat_mom,
                    st->llgrid, st->rep, st->capa, st->ptype};
    for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
    free(st);
}

/* ------------------------------------------------------- material set-up */
/* mcm_sets1.f90:23-51 (model 1 constants) */
static void sets1(double *cm) {
#define C(a, b) cm[((a) - 1) + 5 * ((b) - 1)]
    double prop6 = C(2, 1), prop7 = C(3, 1);
    for (int i = 2; i <= 5; i++)
        for (int j = 1; j <= 5; j++) C(i, j) = 0.0;
    double q1 = C(1, 1) * C(1, 2) / ((1.0 + C(1, 2)) * (1.0 - 2.0 * C(1, 2)));
    double q2 = C(1, 1) * 0.5 / (1.0 + C(1, 2));
    double q3 = q1 + 2.0 * q2;
    C(2, 1) = q3; C(3, 2) = q3; C(4, 3) = q3; C(5, 4) = q2;
    C(2, 2) = q1; C(2, 3) = q1; C(3, 1) = q1; C(3, 3) = q1; C(4, 1) = q1; C(4, 2) = q1;
    C(2, 6) = C(1, 6);
    C(1, 6) = C(1, 1) * C(1, 4) / (C(1, 1) - C(1, 4));
    C(1, 7) = prop6;
    C(2, 7) = prop7;
#undef C
}

/* mcm_sets3.f90:25-43 (model 3 constants) */
static void sets3(double *cm) {
#define C(a, b) cm[((a) - 1) + 5 * ((b) - 1)]
    double prop7 = C(2, 6);
    for (int i = 2; i <= 5; i++)
        for (int j = 1; j <= 6; j++) C(i, j) = 0.0;
    double q1 = C(1, 1) * C(1, 2) / ((1.0 + C(1, 2)) * (1.0 - 2.0 * C(1, 2)));
    double q2 = C(1, 1) * 0.5 / (1.0 + C(1, 2));
    double q3 = q1 + 2.0 * q2;
    C(2, 1) = q3;
    C(5, 4) = q2;
    C(2, 4) = C(1, 1) * C(1, 4) / (C(1, 1) - C(1, 4));
    C(2, 6) = prop7;
#undef C
}

/* mcm_sets10.f90:23-36 (model 10 defaults: pcut -1e30, spall type 1) */
static void sets10(double *cm) {
    double prop[48];
    for (int i = 0; i < 48; i++) prop[i] = cm[i];
    if (prop[3] == 0.0) prop[3] = (double)(-1.e30f);
    if (prop[6] == 0.0) prop[6] = (double)(1.000000001f);
    for (int i = 0; i < 48; i++) cm[i] = prop[i];
    for (int i = 2; i <= 16; i++)
        if (prop[i + 15 - 1] < prop[i + 16 - 1]) cm[16 - 1] = (double)i;
}

/* initel in mcm_sets4.f90:43-94 (implicit double precision): the initial cap
   parameter kappa of the cap model from the initial cap position xint, and
   the tension cut-off fcut, by Newton iteration with the single-precision
   tolerance 1.0e-3.  The reference warns and keeps iterating past maxit; here
   a run that needs more than 100000 iterations is rejected. */
static int initel(double xint, double alpha, double gama, double beta, double theta, double r, double *capint_out,
                  double *fcut_out) {
    const double tol = (double)1.0e-3f, tiny = (double)1.e-15f;
    double capint = 0.0, delcap = 1.e10, delf = 1.e10, fcut = 0.0;
    for (long it = 0;; it++) {
        if (it > 100000) return -1;
        double fe = alpha - gama * exp(-(beta * capint)) + theta * capint;
        double psik = xint - capint - r * fe;
        double err1 = fabs(psik);
        double err2 = fabs(delcap) / (fabs(capint) + tiny);
        if (!(err1 > tol && err2 > tol)) break;
        double fep = beta * gama * exp(-(beta * capint)) + theta;
        double dpsik = -1.0 - r * fep;
        delcap = psik / dpsik;
        capint = capint - delcap;
    }
    for (long it = 0;; it++) {
        if (it > 100000) return -1;
        double fe = alpha - gama * exp(-(beta * fcut)) + theta * fcut;
        double err1 = fabs(fe);
        double err2 = fabs(delf) / (fabs(fcut) + tiny);
        if (!(err1 > tol && err2 > tol)) break;
        double fep = gama * beta * exp(-(beta * fcut)) + theta;
        delf = fe / fep;
        fcut = fcut - delf;
    }
    *capint_out = capint;
    *fcut_out = fcut;
    return 0;
}

/* mcm_sets4.f90:19-41 (model 4, the cap model): lambda + 2G from the bulk and
   shear moduli in cm(5,2), the initial cap parameter in cm(3,3), the tension
   cut-off in cm(3,4).  Only the second tempNi of the reference is used. */
static int sets4(double *cm) {
#define C(a, b) cm[((a) - 1) + 5 * ((b) - 1)]
    double alpha = C(1, 3), theta = C(1, 4), gama = C(1, 5), beta = C(1, 6), r = C(2, 1);
    double xint = C(2, 5);
    double tempE = (9.0 * C(1, 1) * C(1, 2)) / (3.0 * C(1, 1) + C(1, 2));
    double tempNi = (3.0 * C(1, 1) - tempE) / (6.0 * C(1, 1));
    double q1 = tempE * tempNi / ((1.0 + tempNi) * (1.0 - 2.0 * tempNi));
    double q2 = tempE * 0.5 / (1.0 + tempNi);
    double q3 = q1 + 2.0 * q2;
    C(5, 2) = q3;
    double capint, fcut;
    if (initel(xint, alpha, gama, beta, theta, r, &capint, &fcut)) return -1;
    C(3, 3) = capint;
    C(3, 4) = fcut;
    return 0;
#undef C
}

/* ------------------------------------------------------------------ deck */
#define DECK_MAGIC 0x4D434D31
#define CONTACT_TAG 0x434F4E54   /* 'CONT': penalty contact section after the particles */
#define OPTS_TAG 0x4F505453      /* 'OPTS': dynamic relaxation and base acceleration section */

orc_state *orc_load_deck(const char *path, int ncap, int gcap) {
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    int32_t magic, ints[19];
    double reals[9];
    if (fread(&magic, 4, 1, f) != 1 || magic != DECK_MAGIC || fread(ints, 4, 19, f) != 19 ||
        fread(reals, 8, 9, f) != 9) { fclose(f); return NULL; }
    orc_state *st = xcalloc(1, sizeof(orc_state));
    st->np = ints[0]; st->nummat = ints[1]; st->disctype = ints[2]; st->h_opt = ints[3];
    st->tcrit_opt = ints[4]; st->contacttype = ints[5]; st->krtype = ints[6];
    st->massopt = ints[8]; st->init_h_opt = ints[9]; st->init_rhoe = ints[10];
    st->veloc_opt = ints[18];
    int mirr_opt = ints[11];
    const int *flag = &ints[12];
    st->tssf = reals[0]; st->itss = reals[1]; st->endtime = reals[2];
    /* control defaults, mcm_control_v2.f90:193-200 */
    if (st->tssf == 0.0) st->tssf = 0.8;
    if (st->krtype == 0) st->krtype = 1;
    st->ncap = ncap > 0 ? ncap : 160;
    st->gcap = gcap > 0 ? gcap : 160;
    st->pi = acos(-1.0);
    st->mat = xcalloc(st->nummat, sizeof(orc_material));
    st->mat_ke = xcalloc(st->nummat, 8); st->mat_ie = xcalloc(st->nummat, 8);
    st->mat_mom = xcalloc(3 * st->nummat, 8);
    for (int m = 0; m < st->nummat; m++) {
        orc_material *mt = &st->mat[m];
        int32_t mi[3]; double mr[7];
        if (fread(mi, 4, 3, f) != 3 || fread(mr, 8, 7, f) != 7 || fread(mt->strinput, 8, 48, f) != 48 ||
            fread(mt->eosinput, 8, 24, f) != 24) { fclose(f); orc_free(st); return NULL; }
        mt->model = mi[0]; mt->eos = mi[1]; mt->visc_type = mi[2];
        mt->rho = mr[0]; mt->av_q = mr[1]; mt->av_l = mr[2]; mt->mass = mr[3]; mt->h = mr[4];
        mt->rho_min = mr[5]; mt->rho_max = mr[6];
        /* material card defaults, mcm_matin.f90:46-56 */
        if (mt->visc_type == 0) mt->visc_type = 1;
        if (mt->av_l == 0.0) mt->av_l = 0.06;
        if (mt->av_q == 0.0) mt->av_q = 1.5;
        if (mt->rho_min == 0.0) mt->rho_min = 1.0E-10;
        if (mt->rho_max == 0.0) mt->rho_max = 1.0E+10;
        switch (mt->model) {   /* mcm_matin.f90:152-161 */
            case 1: sets1(mt->strinput); break;
            case 3: sets3(mt->strinput); break;
            case 4: {
                /* the cap model's modes are complete for soil (1) and rock (2)
                   only; another ltype leaves the mode of the previous call */
                int lt = (int)mt->strinput[((2) - 1) + 5 * ((6) - 1)];
                if ((lt != 1 && lt != 2) || sets4(mt->strinput)) set_err(st, ORC_ERR_MODEL);
                break;
            }
            case 10: sets10(mt->strinput); break;
            case 9: break;
            default: set_err(st, ORC_ERR_MODEL); break;
        }
        /* EOS V0 defaults, mcm_printm.f90:84-100 */
        if ((mt->eos == 4 || mt->eos == 41) && mt->eosinput[7] == 0.0) mt->eosinput[7] = 1.0;
        if ((mt->eos == 13 || mt->eos == 28) && mt->eosinput[3] == 0.0) mt->eosinput[3] = 1.0;
    }
    alloc_particles(st);
    long n = st->np;
    int32_t *ibuf = xcalloc(n, 4);
    double *pm = xcalloc(n, 8), *ph = xcalloc(n, 8), *pr = xcalloc(n, 8), *pe = xcalloc(n, 8);
    int ok = fread(st->x, 8, 3 * n, f) == (size_t)(3 * n) && fread(st->v, 8, 3 * n, f) == (size_t)(3 * n);
    ok = ok && fread(ibuf, 4, n, f) == (size_t)n;
    for (long i = 0; i < n; i++) st->pmat[i] = ibuf[i] - 1;
    ok = ok && fread(ibuf, 4, n, f) == (size_t)n;
    for (long i = 0; i < n; i++) st->dispbc[i] = ibuf[i];
    ok = ok && fread(pm, 8, n, f) == (size_t)n && fread(ph, 8, n, f) == (size_t)n &&
         fread(pr, 8, n, f) == (size_t)n && fread(pe, 8, n, f) == (size_t)n;
    if (ok && st->contacttype > 0) {
        /* penalty contact cards, mcm_contact_in.f90:22-45 (binary section after the particles) */
        int32_t ct[3];
        ok = fread(ct, 4, 3, f) == 3 && ct[0] == CONTACT_TAG;
        if (ok) { st->ncontmats = ct[1]; st->cont_opt = ct[2]; }
        for (int m = 0; ok && m < st->nummat; m++) {
            double kn[2];
            ok = fread(kn, 8, 2, f) == 2;
            st->mat[m].k_cont = kn[0]; st->mat[m].n_cont = kn[1];
        }
        /* only "every material in contact" sets K and n; otherwise they are unset */
        if (ok && st->ncontmats != -1) set_err(st, ORC_ERR_MODEL);
    }
    if (ok) {
        /* dynamic relaxation (mcm_control_v2.f90:96,223-225) and base acceleration
           (:112,241-243; mcm_baseaccel_in.f90:25-36) */
        int32_t ot[4];
        if (fread(ot, 4, 4, f) == 4 && ot[0] == OPTS_TAG) {
            double ov[4];
            ok = fread(ov, 8, 4, f) == 4;
            st->drelax_scale = ov[0];
            st->drelax = ov[0] > 0.1 && ov[0] < 1.0;
            st->baseaccel = ot[1] == 1 || ot[2] == 1 || ot[3] == 1;
            for (int d = 0; d < 3; d++) st->base_a[d] = ot[1 + d] == 1 ? ov[1 + d] : 0.0;
        }
    }
    fclose(f);
    if (!ok) { free(ibuf); free(pm); free(ph); free(pr); free(pe); orc_free(st); return NULL; }
    for (long i = 0; i < n; i++) {   /* mcm_nodein.f90:118-135 */
        if (st->massopt == 1) st->mass[i] = pm[i];
        if (st->init_h_opt == 1) st->h[i] = ph[i];
        if (st->init_rhoe == 1) { st->rho[i] = pr[i]; st->e[i] = pe[i]; }
    }
    free(ibuf); free(pm); free(ph); free(pr); free(pe);
    /* boundary planes, mcm_inp_symmetry_v2.f90:83-137 */
    st->boundary = 0;
    if (mirr_opt == 1)
        for (int k = 0; k < 6; k++)
            if (flag[k] != 0) st->boundary = 1;
    for (int d = 0; d < 3; d++) {
        st->boundary_code[d][0] = mirr_opt == 1 ? flag[2 * d] : 0;
        st->boundary_code[d][1] = mirr_opt == 1 ? flag[2 * d + 1] : 0;
        st->boundary_x[d][0] = reals[3 + 2 * d];
        st->boundary_x[d][1] = reals[4 + 2 * d];
        st->boundary_type[d] = 1;
        if (st->boundary_code[d][0] > 0) st->boundary_type[d] += 1;
        if (st->boundary_code[d][1] > 0) st->boundary_type[d] += 2;
    }
    return st;
}

/* --------------------------------------------------------------- kernels */
/* mcm_kernel.f90:42-95 (3D cubic B-spline) */
double orc_kernel_w(const double *xi, const double *xj, double havg, double pi) {
    double const1 = 1.0 / (pi * havg * havg * havg);
    double const2 = const1 / 4.0;
    double rad = 0.0;
    for (int n = 0; n < 3; n++) {
        double xij = xi[n] - xj[n];
        rad = rad + xij * xij;
    }
    rad = sqrt(rad);
    double z = rad / havg;
    if (z < 1.0) return const1 * (((0.75 * z - 1.5) * z) * z + 1.0);
    if (z < 2.0) {
        double t = 2.0 - z;
        return const2 * (t * (t * t));
    }
    return 0.0;
}

/* mcm_gradw.f90:41-103 (gradient with respect to xi; xij = xi - xj) */
void orc_kernel_gradw(const double *xi, const double *xj, double havg, double pi, double *dwdx) {
    double const1 = 1.0 / (pi * havg * havg * havg);
    double const2 = const1 * 0.25;
    double xij[3], rad = 0.0;
    for (int n = 0; n < 3; n++) {
        xij[n] = xi[n] - xj[n];
        rad = rad + xij[n] * xij[n];
    }
    rad = sqrt(rad);
    double z = rad / havg, dw;
    if (z < 1.0) dw = const1 * ((2.25 * z - 3.0) * z);
    else if (z < 2.0) { double t = 2.0 - z; dw = -(3.0 * const2 * (t * t)); }
    else dw = 0.0;
    for (int n = 0; n < 3; n++) dwdx[n] = rad == 0.0 ? 0.0 : dw * xij[n] / (rad * havg);
}

/* mcm_math_func.f90:50-79 (3x3 cofactor inverse, identity if |det| < 1e-8) */
void orc_inverse3(const double *m, double *inv) {
#define M(r, c) m[((r) - 1) + 3 * ((c) - 1)]
#define I(r, c) inv[((r) - 1) + 3 * ((c) - 1)]
    double sd11 = M(2, 2) * M(3, 3) - M(2, 3) * M(3, 2);
    double sd12 = M(2, 1) * M(3, 3) - M(2, 3) * M(3, 1);
    double sd13 = M(2, 1) * M(3, 2) - M(2, 2) * M(3, 1);
    double sd21 = M(1, 2) * M(3, 3) - M(1, 3) * M(3, 2);
    double sd22 = M(1, 1) * M(3, 3) - M(1, 3) * M(3, 1);
    double sd23 = M(1, 1) * M(3, 2) - M(1, 2) * M(3, 1);
    double sd31 = M(1, 2) * M(2, 3) - M(1, 3) * M(2, 2);
    double sd32 = M(1, 1) * M(2, 3) - M(1, 3) * M(2, 1);
    double sd33 = M(1, 1) * M(2, 2) - M(1, 2) * M(2, 1);
    double det = M(1, 1) * sd11 - M(1, 2) * sd12 + M(1, 3) * sd13;
    if (fabs(det) < 1.0e-08) {
        for (int k = 0; k < 9; k++) inv[k] = 0.0;
        I(1, 1) = 1.0; I(2, 2) = 1.0; I(3, 3) = 1.0;
        return;
    }
    double invdet = 1.0 / det;
    I(1, 1) = sd11 * invdet; I(1, 2) = -sd21 * invdet; I(1, 3) = sd31 * invdet;
    I(2, 1) = -sd12 * invdet; I(2, 2) = sd22 * invdet; I(2, 3) = -sd32 * invdet;
    I(3, 1) = sd13 * invdet; I(3, 2) = -sd23 * invdet; I(3, 3) = sd33 * invdet;
#undef M
#undef I
}

/* ---------------------------------------------------------- neighbours */
#define LL(st, i, j, k) ((st)->llgrid[((long)(k) * (st)->gridlim[1] + (j)) * (st)->gridlim[0] + (i)])

/* mcm_setuplist.f90:30-135.  Cell indices here are 0-based: cell = int(...) */
static void setuplist(orc_state *st) {
    const double factor = 2.0;
    for (int i = 0; i < 3; i++) {
        st->gridmin[i] = 0.0;
        st->gridlim[i] = 1;
        st->gridsize[i] = 1.0 / (factor * st->hmax);
    }
    for (int i = 0; i < 3; i++) {
        switch (st->boundary_type[i]) {
            case 1:
                st->gridmin[i] = st->coord_maxmin[i][0] - 0.01 * st->hmax;
                st->gridlim[i] = (int)(st->gridsize[i] * (st->coord_maxmin[i][1] - st->gridmin[i])) + 1;
                break;
            case 2:
                st->gridmin[i] = st->boundary_x[i][0];
                st->gridlim[i] = (int)(st->gridsize[i] * (st->coord_maxmin[i][1] - st->gridmin[i])) + 1;
                break;
            case 3:
                st->gridlim[i] = (int)(st->gridsize[i] * (st->boundary_x[i][1] - st->coord_maxmin[i][0])) + 1;
                st->gridmin[i] = st->boundary_x[i][1] - st->gridlim[i] * factor * st->hmax;
                break;
            case 4:
                st->gridmin[i] = st->boundary_x[i][0];
                st->gridlim[i] = (int)(st->gridsize[i] * (st->boundary_x[i][1] - st->boundary_x[i][0]));
                if (st->gridlim[i] < 1) st->gridlim[i] = 1;
                st->gridsize[i] = (double)(float)st->gridlim[i] / (st->boundary_x[i][1] - st->boundary_x[i][0]);
                break;
        }
    }
    for (int i = 0; i < 3; i++) {
        int c0 = st->boundary_code[i][0], c1 = st->boundary_code[i][1];
        if (c0 == 1 || c0 == 2) {
            st->gridlim[i] += 1;
            st->gridmin[i] = st->gridmin[i] - (1.0 / st->gridsize[i]);
        }
        if (c1 == 1 || c1 == 2) st->gridlim[i] += 1;
    }
    long ncell = (long)st->gridlim[0] * st->gridlim[1] * st->gridlim[2];
    if (ncell > st->llgrid_cap) {
        free(st->llgrid);
        st->llgrid = xcalloc(ncell, sizeof(int));
        st->llgrid_cap = ncell;
    }
    memset(st->llgrid, 0, ncell * sizeof(int));
    for (int i = 0; i < st->np; i++) {
        int box[3], in = 1;
        for (int j = 0; j < 3; j++) {
            box[j] = (int)(st->gridsize[j] * (V3(st->x, i, j) - st->gridmin[j])) + 1;
            if (box[j] < 1) in = 0;
            if (box[j] > st->gridlim[j]) in = 0;
        }
        if (in) {
            st->llpointer[i] = LL(st, box[0] - 1, box[1] - 1, box[2] - 1);
            LL(st, box[0] - 1, box[1] - 1, box[2] - 1) = i + 1;
        } else {
            st->llpointer[i] = -1;
        }
    }
}

/* mcm_init_boundary.f90:28-116 (plane multiply/add vectors) */
static void init_boundary(orc_state *st) {
    for (int p = 0; p < 6; p++) {
        int d = p / 2, side = p % 2, code = st->boundary_code[d][side];
        for (int k = 0; k < 3; k++) { st->plane_mult[p][k] = 1.0; st->plane_add[p][k] = 0.0; }
        if (code == 1) {
            st->plane_mult[p][d] = -1.0;
            st->plane_add[p][d] = 2.0 * st->boundary_x[d][side];
        } else if (code == 2) {
            double span = st->boundary_x[d][1] - st->boundary_x[d][0];
            st->plane_add[p][d] = side == 0 ? -span : span;
        }
    }
}

static void create_ghost(orc_state *st, int from_ghost, int id, int ci, int cj, int ck, int plane) {
    int g = st->ngp;
    if (g >= st->ngp_cap) { set_err(st, ORC_ERR_CAPACITY); return; }
    st->ngp++;
    const double *mult = st->plane_mult[plane], *add = st->plane_add[plane];
    if (!from_ghost) {   /* mcm_ghost_setup.f90:330-341 */
        st->g_par[g] = id; st->g_mat[g] = st->pmat[id]; st->g_mass[g] = st->mass[id];
        st->g_h[g] = st->h[id]; st->g_hold[g] = st->hold[id]; st->g_rho[g] = st->rho[id];
        for (int m = 0; m < 3; m++) {
            V3(st->g_x, g, m) = V3(st->x, id, m) * mult[m] + add[m];
            V3(st->g_xzero, g, m) = V3(st->xzero, id, m) * mult[m] + add[m];
            V3(st->g_v, g, m) = V3(st->v, id, m) * mult[m];
        }
    } else {             /* mcm_ghost_setup.f90:398-409 */
        st->g_par[g] = st->g_par[id]; st->g_mat[g] = st->g_mat[id]; st->g_mass[g] = st->g_mass[id];
        st->g_h[g] = st->g_h[id]; st->g_hold[g] = st->g_hold[id]; st->g_rho[g] = st->g_rho[id];
        for (int m = 0; m < 3; m++) {
            V3(st->g_x, g, m) = V3(st->g_x, id, m) * mult[m] + add[m];
            V3(st->g_xzero, g, m) = V3(st->g_xzero, id, m) * mult[m] + add[m];
            V3(st->g_v, g, m) = V3(st->g_v, id, m) * mult[m];
        }
    }
    st->g_ll[g] = LL(st, ci, cj, ck);
    LL(st, ci, cj, ck) = -(g + 1);
}

/* walk one source cell and mirror its contents into the target cell
   (mcm_ghost_setup.f90:49-68 and the five sibling blocks) */
static void mirror_cell(orc_state *st, int si, int sj, int sk, int ti, int tj, int tk, int plane) {
    int id = LL(st, si, sj, sk);
    if (id > 0) {
        while (id != 0) {
            create_ghost(st, 0, id - 1, ti, tj, tk, plane);
            id = st->llpointer[id - 1];
        }
    } else {
        while (id != 0) {
            int gid = abs(id) - 1;
            create_ghost(st, 1, gid, ti, tj, tk, plane);
            id = st->g_ll[gid];
        }
    }
}

/* mcm_ghost_setup.f90:32-289; ghost storage is sized 4N+1024 here, so the
   capacity-retry branch (:269-282) is never taken */
static void ghost_setup(orc_state *st) {
    st->ngp = 0;
    int L0 = st->gridlim[0], L1 = st->gridlim[1], L2 = st->gridlim[2];
    for (int d = 0; d < 3; d++) {
        for (int side = 0; side < 2; side++) {
            int code = st->boundary_code[d][side];
            if (code != 1 && code != 2) continue;
            int lim = st->gridlim[d];
            int src = side == 0 ? (code == 1 ? 2 : lim - 1) : (code == 1 ? lim - 1 : 2);
            int dst = side == 0 ? 1 : lim;
            int plane = 2 * d + side;
            src -= 1; dst -= 1;   /* 0-based */
            if (d == 0) {
                for (int k = 0; k < L2; k++)
                    for (int j = 0; j < L1; j++) mirror_cell(st, src, j, k, dst, j, k, plane);
            } else if (d == 1) {
                for (int k = 0; k < L2; k++)
                    for (int i = 0; i < L0; i++) mirror_cell(st, i, src, k, i, dst, k, plane);
            } else {
                for (int j = 0; j < L1; j++)
                    for (int i = 0; i < L0; i++) mirror_cell(st, i, j, src, i, j, dst, plane);
            }
        }
    }
}

/* mcm_ll_neighbours.f90:44-299 (lists rebuilt from scratch; capacity errors
   are reported, the reference's two-pass regrow is not reproduced) */
static void ll_neighbours(orc_state *st, int ip_flag) {
#pragma omp parallel for schedule(dynamic, 256) if (st->np >= 16384)
    for (int i = 0; i < st->np; i++) {
        st->nnbr[i] = 0; st->ncont[i] = 0; st->g_nnbr[i] = 0; st->g_ncont[i] = 0;
        st->mindist[i] = 1.0e20;
        double mindist_cp = 1.0e20;
        if (st->llpointer[i] > -1) {
            int lo[3], hi[3];
            for (int j = 0; j < 3; j++) {
                int box = (int)(st->gridsize[j] * (V3(st->x, i, j) - st->gridmin[j])) + 1;
                lo[j] = box - 1 > 1 ? box - 1 : 1;
                hi[j] = box + 1 < st->gridlim[j] ? box + 1 : st->gridlim[j];
            }
            int *nl = st->nbr + (long)i * st->ncap, *cl = st->cont + (long)i * st->ncap;
            int *gl = st->gnbr + (long)i * st->gcap, *gcl = st->gcont + (long)i * st->gcap;
            for (int n = lo[2]; n <= hi[2]; n++)
                for (int m = lo[1]; m <= hi[1]; m++)
                    for (int l = lo[0]; l <= hi[0]; l++) {
                        int id = LL(st, l - 1, m - 1, n - 1);
                        if (id > -1) {
                            while (id >= 1) {
                                int j = id - 1;
                                if (j != i) {
                                    double dist_sq = 0.0;
                                    for (int k = 0; k < 3; k++) {
                                        double dd = V3(st->x, j, k) - V3(st->x, i, k);
                                        dist_sq = dist_sq + dd * dd;
                                    }
                                    double h_avg = 2.0 * 0.5 * (st->h[j] + st->h[i]);
                                    double h_sq = h_avg * h_avg;
                                    if (dist_sq < h_sq) {
                                        if (st->pmat[j] == st->pmat[i]) {
                                            st->mindist[i] = fmin_(st->mindist[i], dist_sq);
                                            if (st->nnbr[i] < st->ncap) nl[st->nnbr[i]] = j; else set_err(st, ORC_ERR_CAPACITY);
                                            st->nnbr[i]++;
                                        } else {
                                            mindist_cp = fmin_(mindist_cp, dist_sq);
                                            if (st->ncont[i] < st->ncap) cl[st->ncont[i]] = j; else set_err(st, ORC_ERR_CAPACITY);
                                            st->ncont[i]++;
                                        }
                                    }
                                }
                                id = st->llpointer[j];
                            }
                        } else {
                            while (id <= -1) {
                                int g = -id - 1;
                                double dist_sq = 0.0;
                                for (int k = 0; k < 3; k++) {
                                    double dd = V3(st->g_x, g, k) - V3(st->x, i, k);
                                    dist_sq = dist_sq + dd * dd;
                                }
                                double h_avg = 2.0 * 0.5 * (st->g_h[g] + st->h[i]);
                                double h_sq = h_avg * h_avg;
                                if (dist_sq < h_sq) {
                                    if (st->g_mat[g] == st->pmat[i]) {
                                        st->mindist[i] = fmin_(st->mindist[i], dist_sq);
                                        if (st->g_nnbr[i] < st->gcap) gl[st->g_nnbr[i]] = g; else set_err(st, ORC_ERR_CAPACITY);
                                        st->g_nnbr[i]++;
                                    } else {
                                        mindist_cp = fmin_(mindist_cp, dist_sq);
                                        if (st->g_ncont[i] < st->gcap) gcl[st->g_ncont[i]] = g; else set_err(st, ORC_ERR_CAPACITY);
                                        st->g_ncont[i]++;
                                    }
                                }
                                id = st->g_ll[g];
                            }
                        }
                    }
            if (st->contacttype == 0) {   /* merge contact lists, :239-267 */
                for (int j = 0; j < st->ncont[i]; j++) {
                    if (st->nnbr[i] + j < st->ncap) nl[st->nnbr[i] + j] = cl[j]; else set_err(st, ORC_ERR_CAPACITY);
                }
                st->nnbr[i] += st->ncont[i];
                st->mindist[i] = fmin_(st->mindist[i], mindist_cp);
                if (st->boundary) {
                    for (int j = 0; j < st->g_ncont[i]; j++) {
                        if (st->g_nnbr[i] + j < st->gcap) gl[st->g_nnbr[i] + j] = gcl[j]; else set_err(st, ORC_ERR_CAPACITY);
                    }
                    st->g_nnbr[i] += st->g_ncont[i];
                }
            }
            st->mindist[i] = sqrt(st->mindist[i]);
        }
    }
    if (ip_flag == 1) {   /* append self, :288-299 */
        for (int i = 0; i < st->np; i++) {
            if (st->nnbr[i] < st->ncap) st->nbr[(long)i * st->ncap + st->nnbr[i]] = i; else set_err(st, ORC_ERR_CAPACITY);
            st->nnbr[i]++;
        }
    }
}

/* mcm_neighbours.f90:29-49 and mcm_init_neighb.f90:294-328 */
int orc_neighbours(orc_state *st) {
    setuplist(st);
    if (st->boundary) ghost_setup(st);
    ll_neighbours(st, st->disctype == 1 ? 1 : 0);
    return st->error;
}

/* -------------------------------------------------------------- strain */
static void nbr_strain_info(const orc_state *st, int i, int k, double *xj, double *vj, double *mj,
                            double *rhoj, double *hj) {
    if (k < st->nnbr[i]) {   /* mcm_get_j_strain_info.f90:30-55 */
        int j = st->nbr[(long)i * st->ncap + k];
        for (int d = 0; d < 3; d++) { xj[d] = V3(st->x, j, d); vj[d] = V3(st->v, j, d); }
        *mj = st->mass[j]; *rhoj = st->rho[j]; *hj = st->h[j];
    } else {
        int g = st->gnbr[(long)i * st->gcap + (k - st->nnbr[i])];
        for (int d = 0; d < 3; d++) { xj[d] = V3(st->g_x, g, d); vj[d] = V3(st->g_v, g, d); }
        *mj = st->g_mass[g]; *rhoj = st->g_rho[g]; *hj = st->g_h[g];
    }
}

static void rod_spin_from(orc_state *st, int i, const double gv[3][3]) {
    for (int n = 0; n < 3; n++)
        for (int m = 0; m < 3; m++) {
            T9(st->rod, i, n, m) = 0.5 * (gv[n][m] + gv[m][n]);
            T9(st->spin, i, n, m) = 0.5 * (gv[n][m] - gv[m][n]);
        }
    st->tracerod[i] = T9(st->rod, i, 0, 0) + T9(st->rod, i, 1, 1) + T9(st->rod, i, 2, 2);
}

/* mcm_oldstrain.f90:34-64 */
static void oldstrain(orc_state *st) {
#pragma omp parallel for schedule(dynamic, 256) if (st->np >= 16384)
    for (int i = 0; i < st->np; i++) {
        double gv[3][3] = {{0}};
        double vi[3] = {V3(st->v, i, 0), V3(st->v, i, 1), V3(st->v, i, 2)};
        int nk = st->nnbr[i] + st->g_nnbr[i];
        for (int k = 0; k < nk; k++) {
            double xj[3], vj[3], mj, rhoj, hj, dwdx[3];
            nbr_strain_info(st, i, k, xj, vj, &mj, &rhoj, &hj);
            double volj = mj / rhoj;
            double havg = 0.5 * (st->h[i] + hj);
            orc_kernel_gradw(&V3(st->x, i, 0), xj, havg, st->pi, dwdx);
            for (int n = 0; n < 3; n++)
                for (int m = 0; m < 3; m++) gv[n][m] = gv[n][m] + volj * (vj[n] - vi[n]) * dwdx[m];
        }
        rod_spin_from(st, i, gv);
    }
}

/* corrected kernel gradient, mcm_corr_strain.f90:88 / mcm_corr_acceleration.f90:104 */
static void corr_grad(const double *dwdx, double w, double sum_w, const double *sum_gradw, double *cg) {
    for (int m = 0; m < 3; m++) cg[m] = (dwdx[m] * sum_w - w * sum_gradw[m]) / (sum_w * sum_w);
}

/* mcm_corr_strain.f90:37-146 (three neighbour passes, kernel centred at i) */
static void corr_strain(orc_state *st) {
#pragma omp parallel for schedule(dynamic, 256) if (st->np >= 16384)
    for (int i = 0; i < st->np; i++) {
        const double *xi = &V3(st->x, i, 0);
        int nk = st->nnbr[i] + st->g_nnbr[i];
        double sum_w = 0.0, sum_gradw[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < nk; k++) {
            double xj[3], vj[3], mj, rhoj, hj, dwdx[3];
            nbr_strain_info(st, i, k, xj, vj, &mj, &rhoj, &hj);
            double volj = mj / rhoj, havg = 0.5 * (st->h[i] + hj);
            double w = orc_kernel_w(xi, xj, havg, st->pi);
            orc_kernel_gradw(xi, xj, havg, st->pi, dwdx);
            sum_w = sum_w + volj * w;
            for (int l = 0; l < 3; l++) sum_gradw[l] = sum_gradw[l] + volj * dwdx[l];
        }
        double B[9] = {0}, invB[9];
        for (int k = 0; k < nk; k++) {
            double xj[3], vj[3], mj, rhoj, hj, dwdx[3], cg[3];
            nbr_strain_info(st, i, k, xj, vj, &mj, &rhoj, &hj);
            double volj = mj / rhoj, havg = 0.5 * (st->h[i] + hj);
            double w = orc_kernel_w(xi, xj, havg, st->pi);
            orc_kernel_gradw(xi, xj, havg, st->pi, dwdx);
            corr_grad(dwdx, w, sum_w, sum_gradw, cg);
            for (int m = 0; m < 3; m++)
                for (int n = 0; n < 3; n++) B[n + 3 * m] = B[n + 3 * m] + volj * (xj[n] - xi[n]) * cg[m];
        }
        orc_inverse3(B, invB);
        double f[3][3] = {{0}};
        for (int k = 0; k < nk; k++) {
            double xj[3], vj[3], mj, rhoj, hj, dwdx[3], cg[3], g2[3];
            nbr_strain_info(st, i, k, xj, vj, &mj, &rhoj, &hj);
            double volj = mj / rhoj, havg = 0.5 * (st->h[i] + hj);
            double w = orc_kernel_w(xi, xj, havg, st->pi);
            orc_kernel_gradw(xi, xj, havg, st->pi, dwdx);
            corr_grad(dwdx, w, sum_w, sum_gradw, cg);
            for (int r = 0; r < 3; r++)   /* matmul(invB, cg) */
                g2[r] = invB[r + 0] * cg[0] + invB[r + 3] * cg[1] + invB[r + 6] * cg[2];
            for (int n = 0; n < 3; n++)
                for (int m = 0; m < 3; m++) f[n][m] = f[n][m] + volj * vj[n] * g2[m];
        }
        rod_spin_from(st, i, f);
    }
}

/* mcm_strain.f90:27-60: half-step positions back for real particles, strain,
   then forward again (x is rewritten as (x - t) + t) */
int orc_strain(orc_state *st) {
    for (int i = 0; i < st->np; i++)
        for (int j = 0; j < 3; j++) V3(st->x, i, j) = V3(st->x, i, j) - 0.5 * V3(st->v, i, j) * st->dt;
    if (st->disctype == 0) oldstrain(st); else corr_strain(st);
    for (int i = 0; i < st->np; i++)
        for (int j = 0; j < 3; j++) V3(st->x, i, j) = V3(st->x, i, j) + 0.5 * V3(st->v, i, j) * st->dt;
    return st->error;
}

/* mcm_update_h.f90:25-54 */
int orc_update_h(orc_state *st) {
    if (st->h_opt == 1) {
        st->hmax = 0.0;
        for (int i = 0; i < st->np; i++)
            if (st->p[i] >= st->pcut[i]) {
                st->hold[i] = st->h[i];
                st->h[i] = st->hold[i] * (1.0 + st->tracerod[i] * st->dt / 3.0);
                st->hmax = fmax_(st->hmax, st->h[i]);
            }
    } else if (st->h_opt == 2) {
        st->hmax = 0.0;
        double power = 1.0 / 3.0;
        for (int i = 0; i < st->np; i++)
            if (st->p[i] >= st->pcut[i]) {
                st->hold[i] = st->h[i];
                st->h[i] = st->h0[i] * pow(st->rho0[i] / st->rho[i], power);
                st->hmax = fmax_(st->hmax, st->h[i]);
            }
    }
    return st->error;
}

/* -------------------------------------------------------- constitutive */
/* mcm_rhoupd.f90:24-71 */
static void rhoupd(orc_state *st) {
    const double othird = 1.0 / 3.0;
    for (int i = 0; i < st->np; i++) {
        const orc_material *mt = &st->mat[st->pmat[i]];
        st->rhoold[i] = st->rho[i];
        st->rho[i] = st->rho[i] * (1.0 - st->tracerod[i] * st->dt);
        for (int pass = 0; pass < 2; pass++) {
            if (st->rho[i] > mt->rho_max) {
                st->rho[i] = st->rho[i] / (1.0 - st->tracerod[i] * st->dt);
                for (int d = 0; d < 3; d++) T9(st->rod, i, d, d) = T9(st->rod, i, d, d) - othird * st->tracerod[i];
                st->tracerod[i] = 0.0;
                if (pass == 1) st->h[i] = st->hold[i];
                st->h[i] = mt->h * pow(mt->rho / st->rho[i], 1.0 / 3.0);
            }
        }
    }
}

/* mcm_stress_rot1.f90:22-199, iorder = 3 (exponential map); the qq update
   (:189-199) multiplies a zero matrix and is not restated */
static void stress_rot1(orc_state *st, int i) {
    const double etol = (double)1.0e-3f;
    const double z1 = 1.0 / 3.0, z2 = 2.0 / 15.0, z3 = 17.0 / 315.0, z4 = 62.0 / 2835.0;
    double sa[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) sa[r][c] = T9(st->sigma, i, r, c);
    double dt2 = 0.5 * st->dt;
    double wzz = -(dt2 * T9(st->spin, i, 0, 1));
    double wyy = dt2 * T9(st->spin, i, 0, 2);
    double wxx = -(dt2 * T9(st->spin, i, 1, 2));
    double wxx2 = wxx * wxx, wyy2 = wyy * wyy, wzz2 = wzz * wzz;
    double qm2 = wxx2 + wyy2 + wzz2;
    double qm = sqrt(qm2);
    double dd = tan(qm * 0.5);
    double aa, cc;
    if (qm < etol) {
        aa = qm2 * 0.25;
        cc = 0.5 * (1.0 + aa * (z1 + aa * (z2 + aa * (z3 + aa * z4))));
    } else {
        cc = dd / qm;
    }
    double bb = 2.0 * cc / (1.0 + dd * dd);
    aa = bb * cc;
    double Q[3][3];
    Q[0][0] = 1.0 - aa * (wyy2 + wzz2);
    Q[1][1] = 1.0 - aa * (wxx2 + wzz2);
    Q[2][2] = 1.0 - aa * (wxx2 + wyy2);
    Q[0][1] = aa * wxx * wyy;
    Q[0][2] = aa * wxx * wzz;
    Q[1][2] = aa * wyy * wzz;
    Q[1][0] = Q[0][1]; Q[2][0] = Q[0][2]; Q[2][1] = Q[1][2];
    double bwxx = bb * wxx, bwyy = bb * wyy, bwzz = bb * wzz;
    Q[0][1] = Q[0][1] - bwzz; Q[1][0] = Q[1][0] + bwzz;
    Q[0][2] = Q[0][2] + bwyy; Q[2][0] = Q[2][0] - bwyy;
    Q[1][2] = Q[1][2] - bwxx; Q[2][1] = Q[2][1] + bwxx;
    /* the reference reads sigma (1,2), (2,3) and (3,1): for the cap model,
       whose update writes the upper triangle only, (3,1) is the last
       rotation's, not the model's (1,3) */
    double s11 = sa[0][0], s22 = sa[1][1], s33 = sa[2][2], s12 = sa[0][1], s23 = sa[1][2], s31 = sa[2][0];
    double qs[3][3];
    for (int r = 0; r < 3; r++) {
        qs[r][0] = Q[r][0] * s11 + Q[r][1] * s12 + Q[r][2] * s31;
        qs[r][1] = Q[r][0] * s12 + Q[r][1] * s22 + Q[r][2] * s23;
        qs[r][2] = Q[r][0] * s31 + Q[r][1] * s23 + Q[r][2] * s33;
    }
    double n11 = qs[0][0] * Q[0][0] + qs[0][1] * Q[0][1] + qs[0][2] * Q[0][2];
    double n22 = qs[1][0] * Q[1][0] + qs[1][1] * Q[1][1] + qs[1][2] * Q[1][2];
    double n33 = qs[2][0] * Q[2][0] + qs[2][1] * Q[2][1] + qs[2][2] * Q[2][2];
    double n12 = qs[0][0] * Q[1][0] + qs[0][1] * Q[1][1] + qs[0][2] * Q[1][2];
    double n23 = qs[1][0] * Q[2][0] + qs[1][1] * Q[2][1] + qs[1][2] * Q[2][2];
    double n31 = qs[0][0] * Q[2][0] + qs[0][1] * Q[2][1] + qs[0][2] * Q[2][2];
    T9(st->sigma, i, 0, 0) = n11; T9(st->sigma, i, 1, 1) = n22; T9(st->sigma, i, 2, 2) = n33;
    T9(st->sigma, i, 0, 1) = n12; T9(st->sigma, i, 1, 2) = n23; T9(st->sigma, i, 2, 0) = n31;
    T9(st->sigma, i, 1, 0) = n12; T9(st->sigma, i, 2, 1) = n23; T9(st->sigma, i, 0, 2) = n31;
}

#define SG(r, c) T9(st->sigma, i, (r) - 1, (c) - 1)
#define SD(r, c) T9(st->s, i, (r) - 1, (c) - 1)
#define RD(r, c) T9(st->rod, i, (r) - 1, (c) - 1)
#define QV(r, c) T9(st->q, i, (r) - 1, (c) - 1)
#define AL(r, c) T9(st->alfa, i, (r) - 1, (c) - 1)

/* pre-update energy increment shared by f3dm1/f3dm3 (mcm_f3dm3.f90:54-56) */
static double einc_pre(orc_state *st, int i) {
    return (RD(1, 1) * (SG(1, 1) + QV(1, 1)) + RD(2, 2) * (SG(2, 2) + QV(2, 2)) + RD(3, 3) * (SG(3, 3) + QV(3, 3)) +
            2.0 * (RD(1, 2) * (SG(1, 2) + QV(1, 2)) + RD(1, 3) * (SG(1, 3) + QV(1, 3)) + RD(2, 3) * (SG(2, 3) + QV(2, 3)))) *
           st->dt;
}

/* post-update energy increment (mcm_f3dm3.f90:154-156) */
static double einc_post(orc_state *st, int i) {
    return 0.5 * ((RD(1, 1) * SG(1, 1) + RD(2, 2) * SG(2, 2) + RD(3, 3) * SG(3, 3) +
                   2.0 * (RD(1, 2) * SG(1, 2) + RD(1, 3) * SG(1, 3) + RD(2, 3) * SG(2, 3))) *
                      st->dt +
                  st->einc[i]);
}

/* mcm_f3dm1.f90:27-71 (elastic) */
static void f3dm1(orc_state *st, int i, const double *cmv) {
#define C(a, b) cmv[((a) - 1) + 5 * ((b) - 1)]
    const double othird = 1.0 / 3.0;
    double ym = C(1, 1), pr = C(1, 2), gg = C(5, 4);
    double gdt = st->dt * gg;
    double gd2 = 2.0 * gdt;
    double blk = st->dt * ym / (1.0 - 2.0 * pr);
    double davg = othird * st->tracerod[i];
    double p_incr = blk * davg;
    st->einc[i] = einc_pre(st, i);
    SG(1, 1) = SG(1, 1) + p_incr + gd2 * (RD(1, 1) - davg);
    SG(2, 2) = SG(2, 2) + p_incr + gd2 * (RD(2, 2) - davg);
    SG(3, 3) = SG(3, 3) + p_incr + gd2 * (RD(3, 3) - davg);
    SG(1, 2) = SG(1, 2) + gd2 * RD(1, 2);
    SG(1, 3) = SG(1, 3) + gd2 * RD(1, 3);
    SG(2, 3) = SG(2, 3) + gd2 * RD(2, 3);
    SG(2, 1) = SG(1, 2); SG(3, 1) = SG(1, 3); SG(3, 2) = SG(2, 3);
    st->einc[i] = einc_post(st, i);
    st->p[i] = -(othird * (SG(1, 1) + SG(2, 2) + SG(3, 3)));
#undef C
}

/* mcm_f3dm3.f90:31-156 (elastic-plastic, kinematic/isotropic hardening) */
static void f3dm3(orc_state *st, int i, const double *cmv) {
#define C(a, b) cmv[((a) - 1) + 5 * ((b) - 1)]
    const double othird = 1.0 / 3.0;
    double qh = C(2, 4), qb = C(1, 5), qs = C(1, 3), ym = C(1, 1), pr = C(1, 2), efps_f = C(2, 6);
    double gg = 2.0 * C(5, 4);
    double gdt = st->dt * gg;
    double qbqh = qb * qh;
    int efps_flag = (int)lround(C(1, 6));
    double blk = st->dt * ym / (1.0 - 2.0 * pr);
    double davg = othird * st->tracerod[i];
    double p_incr = blk * davg;
    st->p[i] = st->p[i] - p_incr;
    st->einc[i] = einc_pre(st, i);
    SG(1, 1) = SG(1, 1) + p_incr + gdt * (RD(1, 1) - davg);
    SG(2, 2) = SG(2, 2) + p_incr + gdt * (RD(2, 2) - davg);
    SG(3, 3) = SG(3, 3) + p_incr + gdt * (RD(3, 3) - davg);
    SG(1, 2) = SG(1, 2) + gdt * RD(1, 2);
    SG(1, 3) = SG(1, 3) + gdt * RD(1, 3);
    SG(2, 3) = SG(2, 3) + gdt * RD(2, 3);
    SG(2, 1) = SG(1, 2); SG(3, 1) = SG(1, 3); SG(3, 2) = SG(2, 3);
    double k = qs + qbqh * st->efps[i];
    SD(1, 1) = st->p[i] + SG(1, 1) - AL(1, 1);
    SD(2, 2) = st->p[i] + SG(2, 2) - AL(2, 2);
    SD(3, 3) = st->p[i] + SG(3, 3) - AL(3, 3);
    SD(1, 2) = SG(1, 2) - AL(1, 2);
    SD(1, 3) = SG(1, 3) - AL(1, 3);
    SD(2, 3) = SG(2, 3) - AL(2, 3);
    SD(2, 1) = SD(1, 2); SD(3, 1) = SD(1, 3); SD(3, 2) = SD(2, 3);
    double j2 = SD(2, 3) * SD(2, 3) + SD(1, 3) * SD(1, 3) + SD(1, 2) * SD(1, 2) - SD(1, 1) * SD(2, 2) -
                SD(2, 2) * SD(3, 3) - SD(1, 1) * SD(3, 3);
    double k2 = 3.0 * j2 - k * k;
    double scle = 0.5 * (1.0 + fsign1(k2));
    double fac1 = 1.0 / (1.5 * gg + qh);
    double fac2 = 1.5 * gg;
    double fac3 = (1.0 - qb) * qh;
    double j1 = sqrt(3.0 * fabs(j2)) + 1.0 - scle;
    double defps = scle * fac1 * (j1 - k);
    st->efps[i] = st->efps[i] + defps;
    double des = scle * fac2 * defps / j1;
    double depn = scle * fac3 * defps / j1;
    SG(1, 1) = SG(1, 1) - des * SD(1, 1);
    SG(2, 2) = SG(2, 2) - des * SD(2, 2);
    SG(3, 3) = SG(3, 3) - des * SD(3, 3);
    SG(1, 2) = SG(1, 2) - des * SD(1, 2);
    SG(1, 3) = SG(1, 3) - des * SD(1, 3);
    SG(2, 3) = SG(2, 3) - des * SD(2, 3);
    SG(2, 1) = SG(1, 2); SG(3, 1) = SG(1, 3); SG(3, 2) = SG(2, 3);
    AL(1, 1) = AL(1, 1) + depn * SD(1, 1);
    AL(2, 2) = AL(2, 2) + depn * SD(2, 2);
    AL(3, 3) = AL(3, 3) + depn * SD(3, 3);
    AL(1, 2) = AL(1, 2) + depn * SD(1, 2);
    AL(1, 3) = AL(1, 3) + depn * SD(1, 3);
    AL(2, 3) = AL(2, 3) + depn * SD(2, 3);
    AL(2, 1) = AL(1, 2); AL(3, 1) = AL(1, 3); AL(3, 2) = AL(2, 3);
    if (efps_flag == 1 && st->efps[i] > efps_f) {
        double press = othird * (SG(1, 1) + SG(2, 2) + SG(3, 3));
        SG(1, 1) = press; SG(1, 2) = 0.0; SG(1, 3) = 0.0;
        SG(2, 2) = press; SG(2, 1) = 0.0; SG(2, 3) = 0.0;
        SG(3, 3) = press; SG(3, 1) = 0.0; SG(3, 2) = 0.0;
    }
    st->einc[i] = einc_post(st, i);
#undef C
}

/* mcm_f3dm9.f90:26-48 (viscous fluid).  The reference reads pc/xmu from
   cm(1:2, mat) and never sets s(1,2) (it writes s(1,3) twice); both kept. */
static void f3dm9(orc_state *st, int i, const double *cmv) {
    const double othird = 1.0 / 3.0;
    int matid = st->pmat[i] + 1;
    double xmu = cmv[1 + 5 * (matid - 1)];
    if (xmu == 0.0) {
        for (int k = 0; k < 9; k++) st->s[9 * (long)i + k] = 0.0;
    } else {
        double davg = -(othird * st->tracerod[i]);
        SD(1, 1) = xmu * (RD(1, 1) + davg);
        SD(2, 2) = xmu * (RD(2, 2) + davg);
        SD(3, 3) = xmu * (RD(3, 3) + davg);
        SD(1, 3) = xmu * RD(1, 2);
        SD(1, 3) = xmu * RD(1, 3);
        SD(2, 3) = xmu * RD(2, 3);
        SD(2, 1) = SD(1, 2); SD(3, 1) = SD(1, 3); SD(3, 2) = SD(2, 3);
    }
}

/* mcm_f3dm10.f90:28-157 (elastic-plastic-hydrodynamic) */
static void f3dm10(orc_state *st, int i, const double *cm) {
    const double othird = 1.0 / 3.0;
    double gg = cm[0], qs = cm[1], qh = cm[2], a1 = cm[4], a2 = cm[5];
    int ispall = (int)cm[6];
    double epf = cm[7];
    double g2 = 2.0 * st->dt * gg;
    double davg = othird * st->tracerod[i];
    st->p[i] = -(othird * (SG(1, 1) + SG(2, 2) + SG(3, 3)));
    st->p_cut[i] = st->pcut[i];
    double ak = qs + qh * st->efps[i] + (a1 + a2 * st->p[i]) * fmax_(0.0, st->p[i]);
    SD(1, 1) = SG(1, 1) + st->p[i] + g2 * (RD(1, 1) + davg);
    SD(2, 2) = SG(2, 2) + st->p[i] + g2 * (RD(2, 2) + davg);
    SD(3, 3) = SG(3, 3) + st->p[i] + g2 * (RD(3, 3) + davg);
    SD(1, 2) = SG(1, 2) + 0.5 * g2 * RD(1, 2);
    SD(1, 3) = SG(1, 3) + 0.5 * g2 * RD(1, 3);
    SD(2, 3) = SG(2, 3) + 0.5 * g2 * RD(2, 3);
    if (ispall == 3) {
        double deltap = st->p[i] - st->p_cut[i];
        double scle = 0.5 * (1.0 + fsign1(deltap));
        st->pcut[i] = scle * st->p_cut[i];
        SD(1, 1) = scle * SD(1, 1); SD(2, 2) = scle * SD(2, 2); SD(3, 3) = scle * SD(3, 3);
        SD(1, 2) = scle * SD(1, 2); SD(1, 3) = scle * SD(1, 3); SD(2, 3) = scle * SD(2, 3);
        SD(2, 1) = SD(1, 2); SD(3, 1) = SD(1, 3); SD(3, 2) = SD(2, 3);
        st->p_cut[i] = st->pcut[i];
    }
    double aj2 = 0.5 * (SD(1, 1) * SD(1, 1) + SD(2, 2) * SD(2, 2) + SD(3, 3) * SD(3, 3)) + SD(1, 2) * SD(1, 2) +
                 SD(1, 3) * SD(1, 3) + SD(2, 3) * SD(2, 3);
    double sj2 = sqrt(3.0 * aj2);
    double defps = st->f3dm10_defps;
    if (st->fail[i] == 1.0) {
        double thrg = othird / gg;
        defps = fdim_(sj2, ak) / ((double)1.e-30f + fdim_(1.0, -(thrg * qh)));
        defps = thrg * fmin_(sj2, defps);
        st->efps[i] = st->efps[i] + defps;
    }
    st->f3dm10_defps = defps;
    ak = fmax_(ak + qh * defps, 0.0);
    double seps = (double)1.e-30f * gg;
    double scle = (ak + seps) / (fmax_(ak, sj2) + seps);
    SD(1, 1) = scle * SD(1, 1); SD(2, 2) = scle * SD(2, 2); SD(3, 3) = scle * SD(3, 3);
    SD(1, 2) = scle * SD(1, 2); SD(1, 3) = scle * SD(1, 3); SD(2, 3) = scle * SD(2, 3);
    SD(2, 1) = SD(1, 2); SD(3, 1) = SD(1, 3); SD(3, 2) = SD(2, 3);
    if (epf != 0.0 && st->efps[i] > epf) {
        st->fail[i] = 0.0;
        for (int k = 0; k < 9; k++) st->s[9 * (long)i + k] = 0.0;
    }
    if (ispall == 2) {
        aj2 = 0.5 * (SD(1, 1) * SD(1, 1) + SD(2, 2) * SD(2, 2) + SD(3, 3) * SD(3, 3)) + SD(1, 2) * SD(1, 2) +
              SD(1, 3) * SD(1, 3) + SD(2, 3) * SD(2, 3) + (double)1.e-12f;
        sj2 = SD(1, 1) * SD(2, 3) * SD(2, 3) + SD(2, 2) * SD(1, 3) * SD(1, 3) + SD(3, 3) * SD(1, 2) * SD(1, 2) -
              SD(1, 1) * SD(2, 2) * SD(3, 3) - 2.0 * SD(1, 2) * SD(1, 3) * SD(2, 3);
        double akt = -(sqrt(27.0 / aj2) * sj2 * 0.5 / aj2);
        akt = copysign(fmin_(fabs(akt), 1.0), akt);
        double ywh = acos(akt) * othird;
        sj2 = 2.0 * sqrt(aj2 * othird) * cos(ywh);
        double davg1 = st->p[i] - sj2 - st->p_cut[i];
        scle = 0.5 * (1.0 + fsign1(davg1));
        st->pcut[i] = scle * st->p_cut[i];
        SD(1, 1) = scle * SD(1, 1); SD(2, 2) = scle * SD(2, 2); SD(3, 3) = scle * SD(3, 3);
        SD(1, 2) = scle * SD(1, 2); SD(1, 3) = scle * SD(1, 3); SD(2, 3) = scle * SD(2, 3);
        SD(2, 1) = SD(1, 2); SD(3, 1) = SD(1, 3); SD(3, 2) = SD(2, 3);
        st->p_cut[i] = st->pcut[i];
    }
}

/* ------------------------------------------------------------ cap model
   mcm_f3dm25.f90 (material model 4, the cap model of granular material).
   f3dm25a, feit and fcit have no implicit none: their undeclared locals are
   REAL*4 (a-h, o-z) or INTEGER (i-n), and the declared ones REAL*8.  Every
   operation below is carried out in the wider kind of its two operands and
   every assignment rounds to its target's kind, as Fortran does; float
   literals are the reference's single-precision literals.  Single-precision
   exp is the C library's expf, the routine flang calls for it.
   The reference's Debug configuration compiles with /Qsave, so these locals
   are static and zero before their first assignment: the local capa that
   the yield test reads (:205-206) is never assigned and so is 0; blk and
   davg (:182) are never assigned either, and their product is unused. */
static float fmaxF(float a, float b) { return a > b ? a : b; }
static float fminF(float a, float b) { return a < b ? a : b; }

/* feit, mcm_f3dm25.f90:926-983: failure-surface iterations (mode 10, vector
   option 0); returns 0 when converged, -1 after maxit iterations (the
   reference stops) */
static int feit(double capan, float xj1tr, float xj2dtr, float alpha, float gama, float beta, float theta, float w,
                float dDyna, float r, double bulk, float shear, float *xj1_out, double *capa, double *sig) {
    const float tol = 1.0e-8f;
    const int maxit = 100;
    int iter = 1;
    *capa = capan;
    float xj1 = xj1tr, xj2d = xj2dtr;
    float s1 = (float)(sig[0] - (double)(xj1tr / 3.0f));
    float s2 = (float)(sig[1] - (double)(xj1tr / 3.0f));
    float s3 = (float)(sig[2] - (double)(xj1tr / 3.0f));
    float s4 = (float)sig[3], s5 = (float)sig[4], s6 = (float)sig[5];
    for (;;) {
        float xk = (float)(*capa + (double)r * ((double)alpha - (double)gama * exp(-((double)beta * *capa)) +
                                                (double)theta * *capa));
        float dhdk = (float)((double)(w * dDyna * expf(-(dDyna * xk))) *
                             (1.0 + (double)r * ((double)theta + (double)(gama * beta) * exp(-((double)beta * *capa)))));
        float omega = theta + gama * beta * expf(-(beta * xj1));
        float dcd = (float)((double)shear + 9.0 * bulk * (double)(omega * omega));
        float f = xj2d - (alpha - gama * expf(-(beta * xj1)) + theta * xj1);
        float delcap = -(3.0f * omega * f / (dcd * dhdk));
        float delam = -(dhdk * delcap / (3.0f * omega));
        *capa = *capa + (double)delcap;
        float fac1 = shear / xj2d;
        float fac2 = (float)(3.0 * bulk * (double)omega);
        sig[0] = sig[0] - (double)(delam * (fac1 * s1 - fac2));
        sig[1] = sig[1] - (double)(delam * (fac1 * s2 - fac2));
        sig[2] = sig[2] - (double)(delam * (fac1 * s3 - fac2));
        sig[3] = sig[3] - (double)(delam * fac1 * s4);
        sig[4] = sig[4] - (double)(delam * fac1 * s5);
        sig[5] = sig[5] - (double)(delam * fac1 * s6);
        xj1 = (float)(sig[0] + sig[1] + sig[2]);
        s1 = (float)(sig[0] - (double)(xj1 / 3.0f));
        s2 = (float)(sig[1] - (double)(xj1 / 3.0f));
        s3 = (float)(sig[2] - (double)(xj1 / 3.0f));
        s4 = (float)sig[3]; s5 = (float)sig[4]; s6 = (float)sig[5];
        xj2d = sqrtf(0.5f * (s1 * s1 + s2 * s2 + s3 * s3) + s4 * s4 + s5 * s5 + s6 * s6);
        f = xj2d - (alpha - gama * expf(-(beta * xj1)) + theta * xj1);
        *xj1_out = xj1;
        if (fabsf(f) <= tol) return 0;
        if (iter < maxit) iter = iter + 1;
        else return -1;
    }
}

/* fcit, mcm_f3dm25.f90:985-1037: cap-surface iterations (mode 11, vector
   option 0) */
static int fcit(double capan, float xj1tr, float xj2dtr, float alpha, float gama, float beta, float theta, float w,
                float dDyna, float r, double bulk, float shear, float *xj1_out, double *capa, double *sig) {
    const float tol = 1.0e-8f;
    const int maxit = 100;
    const float unit = 1.0f;
    int iter = 1;
    *capa = capan;
    float xj1 = xj1tr, xj2d = xj2dtr;
    float s1 = (float)(sig[0] - (double)(xj1tr / 3.0f));
    float s2 = (float)(sig[1] - (double)(xj1tr / 3.0f));
    float s3 = (float)(sig[2] - (double)(xj1tr / 3.0f));
    float s4 = (float)sig[3], s5 = (float)sig[4], s6 = (float)sig[5];
    for (;;) {
        float xk = (float)(*capa + (double)r * ((double)alpha - (double)gama * exp(-((double)beta * *capa)) +
                                                (double)theta * *capa));
        float dhdk = (float)((double)(w * dDyna * expf(-(dDyna * xk))) *
                             (1.0 + (double)r * ((double)theta + (double)(gama * beta) * exp(-((double)beta * *capa)))));
        float dfdk;
        if (*capa >= 0.0)
            dfdk = (float)(-2.0 * (fabs((double)xj1 - *capa) / (double)(r * r) +
                                   ((double)alpha - (double)gama * exp(-((double)beta * *capa)) + (double)theta * *capa) *
                                       ((double)theta + (double)(gama * beta) * exp(-((double)beta * *capa)))));
        else
            dfdk = (float)((double)(-2.0f * xk) *
                           (1.0 + (double)r * ((double)theta + (double)(gama * beta) * exp(-((double)beta * *capa)))) /
                           (double)(r * r));
        float elcap = (float)fmax_(*capa, (double)(0.0f * unit));
        float omega = fminF(-2.0f * fabsf(xj1 - elcap) / (r * r), -1.e-15f * unit);
        float t2 = 2.0f * (xj1 - elcap) / (r * r);
        float dcd = (float)((double)(4.0f * shear * (xj2d * xj2d)) + 9.0 * bulk * (double)(t2 * t2));
        float f = xj2d * xj2d + (xj1 - elcap) * (xj1 - elcap) / (r * r) - (xk - elcap) * (xk - elcap) / (r * r);
        float delcap = -f / (dfdk + dcd * dhdk / (3.0f * omega));
        float delam = -(dhdk * delcap / (3.0f * omega));
        *capa = *capa + (double)delcap;
        /* xk again at the new cap: the convergence test below reads it */
        xk = (float)(*capa + (double)r * ((double)alpha - (double)gama * exp(-((double)beta * *capa)) +
                                          (double)theta * *capa));
        float fac1 = 2.0f * shear;
        float fac2 = (float)(6.0 * bulk * (double)fabsf(xj1 - elcap) / (double)(r * r));
        sig[0] = sig[0] - (double)(delam * (fac1 * s1 + fac2));
        sig[1] = sig[1] - (double)(delam * (fac1 * s2 + fac2));
        sig[2] = sig[2] - (double)(delam * (fac1 * s3 + fac2));
        sig[3] = sig[3] - (double)(delam * fac1 * s4);
        sig[4] = sig[4] - (double)(delam * fac1 * s5);
        sig[5] = sig[5] - (double)(delam * fac1 * s6);
        xj1 = (float)(sig[0] + sig[1] + sig[2]);
        s1 = (float)(sig[0] - (double)(xj1 / 3.0f));
        s2 = (float)(sig[1] - (double)(xj1 / 3.0f));
        s3 = (float)(sig[2] - (double)(xj1 / 3.0f));
        s4 = (float)sig[3]; s5 = (float)sig[4]; s6 = (float)sig[5];
        xj2d = sqrtf(0.5f * (s1 * s1 + s2 * s2 + s3 * s3) + s4 * s4 + s5 * s5 + s6 * s6);
        f = xj2d * xj2d + (xj1 - elcap) * (xj1 - elcap) / (r * r) - (xk - elcap) * (xk - elcap) / (r * r);
        *xj1_out = xj1;
        if (fabsf(f) <= tol) {
            if (xj1 < elcap) {
                xj1 = elcap + fabsf(xj1 - elcap);
                sig[0] = (double)(s1 + xj1 / 3.0f);
                sig[1] = (double)(s2 + xj1 / 3.0f);
                sig[2] = (double)(s3 + xj1 / 3.0f);
                *xj1_out = xj1;
            }
            return 0;
        }
        if (iter < maxit) iter = iter + 1;
        else return -1;
    }
}

/* one unrolled failure-surface iteration of f3dm25a's mode 10 with vector
   option 1 (mcm_f3dm25.f90:489-516, repeated at :517-541, :545-570,
   :571-596); the same statements as feit's loop body */
static void cap_fail_step(double *capa, double *sig, float *x1, float *x2d, float s[6], float alpha, float gama,
                          float beta, float theta, float w, float dDyna, float r, double bulk, float shear) {
    float xk = (float)(*capa + (double)r * ((double)alpha - (double)gama * exp(-((double)beta * *capa)) +
                                            (double)theta * *capa));
    float dhdk = (float)((double)(w * dDyna * expf(-(dDyna * xk))) *
                         (1.0 + (double)r * ((double)theta + (double)(gama * beta) * exp(-((double)beta * *capa)))));
    float omega = theta + gama * beta * expf(-(beta * *x1));
    float dcd = (float)((double)shear + 9.0 * bulk * (double)(omega * omega));
    float f = *x2d - (alpha - gama * expf(-(beta * *x1)) + theta * *x1);
    float delcap = -(3.0f * omega * f / (dcd * dhdk));
    float delam = -(dhdk * delcap / (3.0f * omega));
    *capa = *capa + (double)delcap;
    float fac1 = shear / *x2d;
    float fac2 = (float)(3.0 * bulk * (double)omega);
    for (int k = 0; k < 3; k++) sig[k] = sig[k] - (double)(delam * (fac1 * s[k] - fac2));
    for (int k = 3; k < 6; k++) sig[k] = sig[k] - (double)(delam * fac1 * s[k]);
    *x1 = (float)(sig[0] + sig[1] + sig[2]);
    for (int k = 0; k < 3; k++) s[k] = (float)(sig[k] - (double)(*x1 / 3.0f));
    for (int k = 3; k < 6; k++) s[k] = (float)sig[k];
    *x2d = sqrtf(0.5f * (s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) + s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
}

/* one unrolled cap-surface iteration of mode 11 with vector option 1
   (mcm_f3dm25.f90:663-692, repeated at :697-726, :731-760, :765-794); the
   same statements as fcit's loop body */
static void cap_cap_step(double *capa, double *sig, float *x1, float *x2d, float s[6], float *elcap_out, float alpha,
                         float gama, float beta, float theta, float w, float dDyna, float r, double bulk, float shear) {
    const float unit = 1.0f;
    float xk = (float)(*capa + (double)r * ((double)alpha - (double)gama * exp(-((double)beta * *capa)) +
                                            (double)theta * *capa));
    float dhdk = (float)((double)(w * dDyna * expf(-(dDyna * xk))) *
                         (1.0 + (double)r * ((double)theta + (double)(gama * beta) * exp(-((double)beta * *capa)))));
    float dfdk;
    if (*capa >= 0.0)
        dfdk = (float)(-2.0 * (fabs((double)*x1 - *capa) / (double)(r * r) +
                               ((double)alpha - (double)gama * exp(-((double)beta * *capa)) + (double)theta * *capa) *
                                   ((double)theta + (double)(gama * beta) * exp(-((double)beta * *capa)))));
    else
        dfdk = (float)((double)(-2.0f * xk) *
                       (1.0 + (double)r * ((double)theta + (double)(gama * beta) * exp(-((double)beta * *capa)))) /
                       (double)(r * r));
    float elcap = (float)fmax_(*capa, (double)(0.0f * unit));
    float omega = fminF(-2.0f * fabsf(*x1 - elcap) / (r * r), -1.e-15f * unit);
    float t2 = 2.0f * (*x1 - elcap) / (r * r);
    float dcd = (float)((double)(4.0f * shear * (*x2d * *x2d)) + 9.0 * bulk * (double)(t2 * t2));
    float f = *x2d * *x2d + (*x1 - elcap) * (*x1 - elcap) / (r * r) - (xk - elcap) * (xk - elcap) / (r * r);
    float delcap = -f / (dfdk + dcd * dhdk / (3.0f * omega));
    float delam = -(dhdk * delcap / (3.0f * omega));
    *capa = *capa + (double)delcap;
    float fac1 = 2.0f * shear;
    float fac2 = (float)(6.0 * bulk * (double)fabsf(*x1 - elcap) / (double)(r * r));
    for (int k = 0; k < 3; k++) sig[k] = sig[k] - (double)(delam * (fac1 * s[k] + fac2));
    for (int k = 3; k < 6; k++) sig[k] = sig[k] - (double)(delam * fac1 * s[k]);
    *x1 = (float)(sig[0] + sig[1] + sig[2]);
    for (int k = 0; k < 3; k++) s[k] = (float)(sig[k] - (double)(*x1 / 3.0f));
    for (int k = 3; k < 6; k++) s[k] = (float)sig[k];
    *x2d = sqrtf(0.5f * (s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) + s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    *elcap_out = elcap;
}

/* f3dm25a, mcm_f3dm25.f90:34-838: trial stress, mode selection (:229-261),
   the update of the mode (:263-797) and the stress, rate of deformation and
   pressure written back (:801-831).  Stresses are negated on entry
   (compression positive) and on exit; only the upper triangle of sigma is
   written, so its lower triangle keeps the rotated stress of stress_rot1. */
static void f3dm25(orc_state *st, int i, const double *cmv) {
#define C(a, b) cmv[((a) - 1) + 5 * ((b) - 1)]
    const float tol = 1.0e-7f;
    const double othird = 1.0 / 3.0;
    const float third = 1.0f / 3.0f;
    const float unit = 1.0f, geop = 0.0f;
    const float capa_l = 0.0f;   /* the never-assigned local capa */
    double *capa = &st->capa[i];
    double bulk = C(1, 1);
    float shear = (float)C(1, 2), alpha = (float)C(1, 3), theta = (float)C(1, 4), gama = (float)C(1, 5);
    float beta = (float)C(1, 6), r = (float)C(2, 1), dDyna = (float)C(2, 2), w = (float)C(2, 3);
    float falfac = (float)C(4, 1);
    alpha = alpha - falfac;
    float z = (float)C(2, 4);
    int ivec = 1;
    if (C(3, 5) != 0.0) ivec = 0;
    int ltype = (int)C(2, 6);
    float tcut = (float)C(3, 1), fcut = (float)C(3, 4);
    float t = fmaxF(fcut, tcut + 3.0f * geop);
    float fet = alpha - gama * expf(-(beta * t)) + theta * t;
    float fept = gama * beta * expf(-(beta * t)) + theta;
    *capa = fmax_(*capa, C(3, 3));
    double sig[6] = {-SG(1, 1), -SG(2, 2), -SG(3, 3), -SG(1, 2), -SG(1, 3), -SG(2, 3)};
    double dd1 = -RD(1, 1), dd2 = -RD(2, 2), dd3 = -RD(3, 3);
    double dd4 = -(RD(1, 2) / 2.0), dd5 = -(RD(1, 3) / 2.0), dd6 = -(RD(2, 3) / 2.0);
    double dd = (dd1 + dd2 + dd3);
    double davgm = (double)(1.0f / 3.0f) * dd;
    float dt1 = (float)st->dt;
    double p = 3.0 * bulk * (double)dt1 * davgm;
    double gdt = st->dt * (double)shear;
    double gd2 = 2.0 * gdt;
    float gtr11 = (float)(sig[0] + p + gd2 * (dd1 - davgm));
    float gtr22 = (float)(sig[1] + p + gd2 * (dd2 - davgm));
    float gtr33 = (float)(sig[2] + p + gd2 * (dd3 - davgm));
    float gtr12 = (float)(sig[3] + gd2 * dd4);
    float gtr23 = (float)(sig[4] + gd2 * dd5);
    float gtr31 = (float)(sig[5] + gd2 * dd6);
    float x1tr = gtr11 + gtr22 + gtr33;
    float s1 = gtr11 - (1.0f / 3.0f) * x1tr, s2 = gtr22 - (1.0f / 3.0f) * x1tr, s3 = gtr33 - (1.0f / 3.0f) * x1tr;
    float sdtr11 = s1, sdtr22 = s2, sdtr33 = s3;
    float x2dtr = 0.5f * (s1 * s1 + s2 * s2 + s3 * s3 + 2.0f * (gtr12 * gtr12 + gtr23 * gtr23 + gtr31 * gtr31));
    x2dtr = sqrtf(x2dtr);
    float fe = alpha - gama * expf(-(beta * capa_l)) + theta * capa_l;
    float xkn = capa_l + r * fe;
    float elcapn = (float)fmax_(*capa, (double)(0.0f * unit));
    float tmises = fabsf(xkn - elcapn) / r;
    fe = alpha - gama * expf(-(beta * x1tr)) + theta * x1tr;
    float f1tr = x2dtr - fminF(fe, tmises);
    float tcrit = (float)((double)t - 9.0 * bulk / (double)shear * (double)(x2dtr - fet) * (double)fept);
    fe = alpha - gama * expf(-(beta * elcapn)) + theta * elcapn;
    float fep = beta * gama * expf(-(beta * elcapn)) + theta;
    float fc2tr = ((xkn - elcapn) * (xkn - elcapn) - (x1tr - elcapn) * (x1tr - elcapn)) / (r * r);
    float f2tr = x2dtr * x2dtr - fc2tr;
    float x1crt = (float)((double)elcapn - (double)((x2dtr - fe) * 9.0f) * bulk * (double)fep / (double)shear);
    fe = (float)((double)alpha - (double)gama * exp(-((double)beta * *capa)) + (double)theta * *capa);
    float capal = (float)fmin_((double)(0.0f * unit), *capa - (double)(tol * fe));
    fe = alpha - gama * expf(-(beta * capal)) + theta * capal;
    float xk = capal + r * fe;
    float evpn = w * (1.0f - expf(dDyna * (z - xkn)));
    float devpb = w * (1.0f - expf(dDyna * (z - xk))) - evpn;
    double capan = *capa;
    int imode = st->cap_imode;   /* a saved local: the last call's mode if no case below matches */
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
    st->cap_imode = imode;
    st->ptype[i] = imode;
    float x1 = 0.0f, x2d = 0.0f, ratio;
    double pp;
    const float s_tr[3] = {s1, s2, s3};
    const float g_sh[3] = {gtr12, gtr23, gtr31};
    switch (imode) {
        case 1: case 2: case 3:
            /* :263-315 elastic, and the tension cut-off without a cap change */
            x1 = imode == 1 ? x1tr : t;
            if (imode == 1) *capa = capan;
            else if (imode == 2) *capa = 0.0;
            else *capa = capan;
            pp = (double)(third * x1 + geop);
            for (int k = 0; k < 3; k++) sig[k] = (double)s_tr[k] + pp;
            for (int k = 0; k < 3; k++) sig[3 + k] = (double)g_sh[k];
            break;
        case 4: case 5:
            /* :320-354 tension corner */
            x1 = t;
            x2d = fet;
            *capa = imode == 4 ? 0.0 : capan;
            pp = (double)(third * x1 + geop);
            ratio = x2d / x2dtr;
            for (int k = 0; k < 3; k++) sig[k] = (double)(s_tr[k] * ratio) + pp;
            for (int k = 0; k < 3; k++) sig[3 + k] = (double)(g_sh[k] * ratio);
            break;
        case 6: case 7: {
            /* :357-398 tension corner / cut-off with the cap moving */
            x1 = t;
            x2d = imode == 6 ? fet : x2dtr;
            float x1old = (float)(sig[0] + sig[1] + sig[2]);
            float devp = (float)(dd - (double)(x1 - x1old) / (3.0 * bulk));
            float xhat = fminF(devpb, devp);
            *capa = *capa + (double)devp * ((double)capal - capan) / (double)xhat;
            *capa = fmax_(*capa, (double)(0.0f * unit));
            pp = (double)(third * x1 + geop);
            if (imode == 6) {
                ratio = x2d / x2dtr;
                for (int k = 0; k < 3; k++) sig[k] = (double)(s_tr[k] * ratio) + pp;
                for (int k = 0; k < 3; k++) sig[3 + k] = (double)(g_sh[k] * ratio);
            } else {
                for (int k = 0; k < 3; k++) sig[k] = (double)s_tr[k] + pp;
                for (int k = 0; k < 3; k++) sig[3 + k] = (double)g_sh[k];
            }
            break;
        }
        case 8: case 9:
            /* :405-442 failure surface without cap motion / at the cap corner */
            if (imode == 8) {
                x1 = x1tr;
                fe = alpha - gama * expf(-(beta * x1)) + theta * x1;
                x2d = fminF(fe, tmises);
            } else {
                x1 = elcapn;
                x2d = alpha - gama * expf(-(beta * x1)) + theta * x1;
            }
            *capa = capan;
            ratio = x2d / x2dtr;
            pp = (double)(third * x1 + geop);
            for (int k = 0; k < 3; k++) sig[k] = (double)(s_tr[k] * ratio) + pp;
            for (int k = 0; k < 3; k++) sig[3 + k] = (double)(g_sh[k] * ratio);
            break;
        case 10: {
            /* :445-624 failure surface with the cap moving */
            const float gt[6] = {gtr11, gtr22, gtr33, gtr12, gtr23, gtr31};
            for (int k = 0; k < 6; k++) sig[k] = (double)gt[k];
            if (ivec == 0) {
                if (feit(capan, x1tr, x2dtr, alpha, gama, beta, theta, w, dDyna, r, bulk, shear, &x1, capa, sig)) {
                    set_err(st, ORC_ERR_MODEL);
                    return;
                }
            } else {
                *capa = capan;
                x1 = x1tr;
                x2d = x2dtr;
                float sv[6];
                for (int k = 0; k < 3; k++) sv[k] = (float)(sig[k] - (double)(x1tr / 3.0f));
                for (int k = 3; k < 6; k++) sv[k] = (float)sig[k];
                for (int it = 0; it < 4; it++)
                    cap_fail_step(capa, sig, &x1, &x2d, sv, alpha, gama, beta, theta, w, dDyna, r, bulk, shear);
            }
            if (x1 > elcapn) x1 = elcapn;
            if (ltype == 2 || capan == 0.0) *capa = capan;
            if ((double)x1 > *capa) *capa = (double)x1;
            if (ltype == 1 && capan >= 0.0) *capa = fmax_(*capa, (double)(0.0f * unit));
            fe = alpha - gama * expf(-(beta * x1)) + theta * x1;
            x2d = fminF(fe, tmises);
            ratio = x2d / x2dtr;
            pp = (double)(third * x1 + geop);
            /* ivec 0 scales the trial deviator s1..s3, ivec 1 its saved copy: the same values */
            const float sd[3] = {sdtr11, sdtr22, sdtr33};
            for (int k = 0; k < 3; k++) sig[k] = (double)((ivec == 0 ? s_tr[k] : sd[k]) * ratio) + pp;
            for (int k = 0; k < 3; k++) sig[3 + k] = (double)(g_sh[k] * ratio);
            break;
        }
        case 11: {
            /* :626-797 cap surface */
            const float gt[6] = {gtr11, gtr22, gtr33, gtr12, gtr23, gtr31};
            for (int k = 0; k < 6; k++) sig[k] = (double)gt[k];
            if (ivec == 0) {
                if (fcit(capan, x1tr, x2dtr, alpha, gama, beta, theta, w, dDyna, r, bulk, shear, &x1, capa, sig)) {
                    set_err(st, ORC_ERR_MODEL);
                    return;
                }
            } else {
                *capa = capan;
                x1 = x1tr;
                x2d = x2dtr;
                float sv[6], elcap = 0.0f;
                for (int k = 0; k < 3; k++) sv[k] = (float)(sig[k] - (double)(x1tr / 3.0f));
                for (int k = 3; k < 6; k++) sv[k] = (float)sig[k];
                for (int it = 0; it < 4; it++)
                    cap_cap_step(capa, sig, &x1, &x2d, sv, &elcap, alpha, gama, beta, theta, w, dDyna, r, bulk, shear);
                if (x1 < elcap) {
                    x1 = elcap + fabsf(x1 - elcap);
                    for (int k = 0; k < 3; k++) sig[k] = (double)(sv[k] + x1 / 3.0f);
                }
            }
            break;
        }
        default: break;   /* mode 0: the stress stays the rotated one */
    }
    SG(1, 1) = -sig[0]; SG(2, 2) = -sig[1]; SG(3, 3) = -sig[2];
    SG(1, 2) = -sig[3]; SG(1, 3) = -sig[4]; SG(2, 3) = -sig[5];
    RD(1, 1) = -dd1; RD(2, 2) = -dd2; RD(3, 3) = -dd3;
    RD(1, 2) = -(dd4 * 2.0); RD(1, 3) = -(dd5 * 2.0); RD(2, 3) = -(dd6 * 2.0);
    st->p[i] = -(othird * (SG(1, 1) + SG(2, 2) + SG(3, 3)));
#undef C
}

/* mcm_nodalav.f90:32-47 via mcm_aviscosity.f90:28-39 */
static void aviscosity(orc_state *st, int i) {
    const orc_material *mt = &st->mat[st->pmat[i]];
    if (mt->visc_type != 1) return;
    for (int k = 0; k < 9; k++) st->q[9 * (long)i + k] = 0.0;
    if (st->tracerod[i] < 0.0) {
        double ht = st->h[i] * st->tracerod[i];
        double q = mt->av_l * st->rho[i] * st->c[i] * st->h[i] * fabs(st->tracerod[i]) + mt->av_q * st->rho[i] * (ht * ht);
        QV(1, 1) = q; QV(2, 2) = q; QV(3, 3) = q;
    }
}

/* mcm_lieupd.f90:26-33: dot_product of q and rod in column-major order */
static void lieupd(orc_state *st, int i) {
    double qpm = 0.0;
    for (int k = 0; k < 9; k++) qpm = qpm + st->q[9 * (long)i + k] * st->rod[9 * (long)i + k];
    st->e[i] = st->e[i] + (st->einc[i] - st->dt * qpm) * st->mass[i] / st->rho[i];
}

/* mcm_hieupd.f90:28-64 (trial energy) */
static void hieupd(orc_state *st, int i) {
    const double othird = 1.0 / 3.0;
    double pold = -(othird * (SG(1, 1) + SG(2, 2) + SG(3, 3)));
    double trace_q = othird * (QV(1, 1) + QV(2, 2) + QV(3, 3));
    double volold = st->mass[i] / st->rhoold[i];
    double volnew = st->mass[i] / st->rho[i];
    double de[3][3];
    for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) {
            double sg = T9(st->sigma, i, k, l), sd = T9(st->s, i, k, l), qv = T9(st->q, i, k, l), rd = T9(st->rod, i, k, l);
            if (k == l) de[k][l] = rd * (0.5 * (sg + pold + sd) - qv + trace_q);
            else de[k][l] = rd * (0.5 * (sg + sd) - qv);
        }
    double vavg = 0.5 * (volnew + volold);
    double eincr = 0.0;
    for (int l = 0; l < 3; l++)
        for (int k = 0; k < 3; k++) eincr = eincr + de[k][l];
    eincr = vavg * eincr;
    double dvol = volnew - volold;
    double mean_incr = dvol * (0.5 * pold + trace_q);
    st->etry[i] = st->e[i] + st->dt * eincr - mean_incr;
    for (int k = 0; k < 9; k++) st->qold[9 * (long)i + k] = st->q[9 * (long)i + k];
}

/* sound speed: mcm_sueos.f90:24-46 -> mcm_eos{1,4,13,28,41}c.f90 */
static void sueos(orc_state *st, int i) {
    const orc_material *mt = &st->mat[st->pmat[i]];
    const double *ei = mt->eosinput;
    double rho = st->rho[i], rho0 = st->rho0[i], mass = st->mass[i];
    switch (mt->eos) {
        case 1: {   /* mcm_eos1c.f90:32-81 */
            double c1 = ei[1], c2 = ei[2], c3 = ei[3], c4 = ei[4], c5 = ei[5], c6 = ei[6];
            double mu = rho / rho0 - 1.0;
            double mu_t = fmax_(mu, 0.0);
            double e1 = 0.0;
            double dmudrho = 1.0 / rho0;
            double termone = c1 * dmudrho + (c2 * (dmudrho * dmudrho)) + (3.0 * c3 * dmudrho) * mu * mu;
            double termtwo = (c5 * dmudrho + c6 * dmudrho * dmudrho) * e1;
            double dpdrho = termone + termtwo;
            double dpde = c4 + c5 * mu + c6 * (mu_t * mu_t);
            double mr = mass / rho;
            double c_sq = ((4.0 / 3.0) * mt->g / rho0 + dpdrho + st->p[i] / rho0 * (mr * mr) * dpde);
            st->c[i] = c_sq > 0.0 ? sqrt(c_sq) : sqrt(fabs(c_sq));
            break;
        }
        case 4: {   /* mcm_eos4c.f90:33-89 */
            double c1 = ei[0], s1 = ei[1], s2 = ei[2], s3 = ei[3], gamma = ei[4], a1 = ei[5];
            double e1 = st->e[i] * rho0 / mass;
            double mu = rho / rho0 - 1.0;
            double dpdrho;
            if (mu > 0.0) {
                double termone = mu * (1.0 - (gamma / 2.0));
                double termtwo = (a1 / 2.0) * mu * mu;
                double eosnum = mu * rho0 * (c1 * c1) * (1.0 + termone - termtwo);
                double mp1 = mu + 1.0;
                double mu11 = mu / mp1;
                double mu21 = (mu * mu) / mp1;
                double mu22 = (mu * mu) / (mp1 * mp1);
                double mu32 = (mu * mu * mu) / (mp1 * mp1);
                double mu33 = (mu * mu * mu) / (mp1 * mp1 * mp1);
                double eosden = 1.0 - (s1 - 1.0) * mu - s2 * mu21 - s3 * mu32;
                double termthree = rho0 * (c1 * c1) * (1.0 + 2.0 * termone - 3.0 * termtwo);
                double termfour = 1.0 - s1 - 2.0 * s2 * mu11 + s2 * mu22 - 3.0 * s3 * mu22 + 2.0 * s3 * mu33;
                double ed2 = eosden * eosden;
                double ed3 = eosden * ed2;
                dpdrho = (termthree / ed2 - 2.0 * eosnum * termfour / ed3) / rho0 + a1 * e1 / rho0;
            } else {
                dpdrho = c1 * c1 + a1 * e1 / rho0;
            }
            double dpde = gamma + a1 * mu;
            double mr = mass / rho;
            double c_sq = ((4.0 / 3.0) * mt->g / rho0 + dpdrho + st->p[i] / rho0 * (mr * mr) * dpde);
            st->c[i] = c_sq > 0.0 ? sqrt(c_sq) : sqrt(fabs(c_sq));
            break;
        }
        case 13: {  /* mcm_eos13c.f90:29-37 */
            double gamma = ei[0];
            double e1 = st->e[i] / mass;
            st->c[i] = sqrt((gamma - 1.0) * gamma * e1);
            break;
        }
        case 28: {  /* mcm_eos28c.f90:28-33 */
            st->c[i] = sqrt(ei[0] * ei[1] / rho0);
            break;
        }
        case 41: {  /* mcm_eos41c.f90:31-85 */
            double cc = ei[0], sl = ei[1], gamma0 = ei[4];
            double e1 = st->e[i] / mass;
            double ao = rho0 * (cc * cc);
            double mu = rho / rho0 - 1.0;
            double gamma = gamma0 * rho0 / rho;
            double bo = ao * (1.0 + 2.0 * (sl - 1.0));
            double slm = sl - 1.0;
            double co = ao * (2.0 * (sl - 1.0) + 3.0 * (slm * slm));
            double f = gamma * mu;
            double a1 = ao * (1.0 - f);
            double b1 = bo * (2.0 - 1.5 * f);
            double c1 = co * (3.0 - 2.0 * f);
            double dpdrho, c2;
            if (mu > 0.0) dpdrho = (a1 + b1 * mu + c1 * (mu * mu)) / rho0 + gamma * e1;
            else dpdrho = ao * (1.0 - f) / rho0 + gamma * e1;
            if (st->p[i] == st->pcut[i]) {
                c2 = cc * cc;
            } else {
                double dpdi = gamma * rho;
                double didrho = st->p[i] / (rho * rho);
                dpdrho = dpdrho + dpdi * didrho;
                c2 = dpdrho;
            }
            if (c2 < 0.0) set_err(st, ORC_ERR_MODEL);
            st->c[i] = sqrt(c2);
            break;
        }
        default: break;
    }
}

/* pressure/energy: mcm_eqos.f90:24-55 -> mcm_eqos{1,4,13,28,41}.f90 */
static void eqos(orc_state *st, int i) {
    const orc_material *mt = &st->mat[st->pmat[i]];
    const double *ei = mt->eosinput;
    double rho = st->rho[i], rho0 = st->rho0[i], mass = st->mass[i];
    switch (mt->eos) {
        case 1: {   /* mcm_eqos1.f90:32-75 */
            double c0 = ei[0], c1 = ei[1], c2 = ei[2], c3 = ei[3], c4 = ei[4], c5 = ei[5], c6 = ei[6];
            double mu = rho / rho0 - 1.0;
            double mu_t = fmax_(mu, 0.0);
            double dvol = mass / rho - mass / st->rhoold[i];
            double vol0 = mass / rho0;
            double e1try = st->etry[i] / vol0;
            double termone = c0 + (c1 + c3 * (mu * mu)) * mu + c2 * (mu_t * mu_t);
            double termtwo = c4 + c5 * mu + c6 * (mu_t * mu_t);
            st->p[i] = (termone + termtwo * e1try) / (1.0 + termtwo * (dvol / vol0));
            if (st->p[i] < st->pcut[i]) {
                st->p[i] = st->pcut[i];
                st->rho[i] = st->rho[i] / (1.0 - st->tracerod[i] * st->dt);
            } else {
                st->e[i] = st->etry[i] - 0.5 * dvol * st->p[i];
            }
            break;
        }
        case 4: {   /* mcm_eqos4.f90:30-82 */
            double c1 = ei[0], s1 = ei[1], s2 = ei[2], s3 = ei[3], gamma = ei[4], a1 = ei[5];
            double mu = rho / rho0 - 1.0;
            double dvol = mass / rho - mass / st->rhoold[i];
            double vol0 = mass / rho0;
            double e1try = st->etry[i] / vol0;
            if (mu > 0.0) {
                double termone = mu * (1.0 - (gamma / 2.0));
                double termtwo = (a1 / 2.0) * mu * mu;
                double eosnum = mu * rho0 * (c1 * c1) * (1.0 + termone - termtwo);
                double mp1 = mu + 1.0;
                double mu21 = (mu * mu) / mp1;
                double mu32 = (mu * mu * mu) / (mp1 * mp1);
                double eosden = 1.0 - (s1 - 1.0) * mu - s2 * mu21 - s3 * mu32;
                st->p[i] = (eosnum / (eosden * eosden) + (gamma + a1 * mu) * e1try) /
                           (1.0 + 0.5 * (gamma + a1 * mu) * dvol / vol0);
            } else {
                st->p[i] = (rho0 * (c1 * c1) * mu + (gamma + a1 * mu) * e1try) /
                           (1.0 + 0.5 * (gamma + a1 * mu) * dvol / vol0);
            }
            if (st->p[i] < st->pcut[i]) {
                st->p[i] = st->pcut[i];
                st->rho[i] = st->rho[i] / (1.0 - st->tracerod[i] * st->dt);
            } else {
                st->e[i] = st->etry[i] - 0.5 * dvol * st->p[i];
            }
            break;
        }
        case 13: {  /* mcm_eqos13.f90:30-46 */
            double gamma_m1 = ei[0] - 1.0;
            double e1try = st->etry[i] / mass;
            double dvol = mass / rho - mass / st->rhoold[i];
            st->p[i] = (rho * gamma_m1 * e1try) / (1.0 + 0.5 * rho * gamma_m1 * dvol / mass);
            st->e[i] = st->etry[i] - 0.5 * dvol * st->p[i];
            break;
        }
        case 28: {  /* mcm_eqos28.f90:30-45 */
            double b = ei[0], gamma = ei[1];
            double dvol = mass / rho - mass / st->rhoold[i];
            st->p[i] = b * (pow(rho / rho0, gamma) - 1.0);
            st->e[i] = st->etry[i] - 0.5 * dvol * st->p[i];
            break;
        }
        case 41: {  /* mcm_eqos41.f90:32-102 */
            double cc = ei[0], sl = ei[1], gama0 = ei[4];
            double etry1 = st->etry[i] / mass;
            double ao = rho0 * (cc * cc);
            double mu = rho / rho0 - 1.0;
            double dvol = mass / rho - mass / st->rhoold[i];
            double vol0 = mass / rho0;
            double gama = gama0 * rho0 / rho;
            double slm = sl - 1.0;
            double bo = ao * (1.0 + 2.0 * (sl - 1.0));
            double co = ao * (2.0 * (sl - 1.0) + 3.0 * (slm * slm));
            double f = gama * mu;
            double ph;
            if (mu > 0.0) ph = ao * mu + bo * (mu * mu) + co * (mu * (mu * mu));
            else ph = ao * mu;
            double pn = (ph * (1.0 - 0.5 * f) + rho * gama * etry1) / (1.0 + 0.5 * rho * gama * dvol / vol0);
            if (pn < st->pcut[i]) {
                pn = st->pcut[i];
                st->rho[i] = st->rho[i] / (1.0 - st->tracerod[i] * st->dt);
            } else {
                st->e[i] = etry1 * mass - 0.5 * dvol * pn;
            }
            st->p[i] = pn;
            break;
        }
        default: break;
    }
}

/* mcm_stress_up.f90:28-36 */
static void stress_up(orc_state *st, int i) {
    SG(1, 1) = SD(1, 1) - st->p[i];
    SG(2, 2) = SD(2, 2) - st->p[i];
    SG(3, 3) = SD(3, 3) - st->p[i];
    SG(1, 2) = SD(1, 2); SG(1, 3) = SD(1, 3); SG(2, 3) = SD(2, 3);
    SG(2, 1) = SG(1, 2); SG(3, 1) = SG(1, 3); SG(3, 2) = SG(2, 3);
}

/* mcm_calccritts.f90:32-87 */
static void calccritts(orc_state *st) {
    st->critts = 1.0e20;
    for (int i = 0; i < st->np; i++) {
        const orc_material *mt = &st->mat[st->pmat[i]];
        if (st->tcrit_opt == 0) {
            double mr = st->mass[i] / st->rho[i];
            double veae = pow(mr, 1.0 / 3.0);
            double ae = pow(mr, 2.0 / 3.0);
            double qu = (mt->av_l * st->c[i] + mt->av_q * st->mass[i] / st->rho[i] * st->tracerod[i]) / ae;
            st->critts_p[i] = veae / (qu + sqrt(qu * qu + st->c[i] * st->c[i]));
        } else if (st->tcrit_opt == 1) {
            st->critts_p[i] = st->h[i] / (st->c[i] + st->vabs[i]);
        } else if (st->tcrit_opt == 2) {
            st->critts_p[i] = st->mindist[i] / (st->c[i] + st->vabs[i]);
        }
        if (st->critts_p[i] < st->critts) st->critts = st->critts_p[i];
    }
}

/* mcm_constitutive.f90:26-87 */
int orc_constitutive(orc_state *st) {
    rhoupd(st);
    for (int i = 0; i < st->np; i++) {
        const orc_material *mt = &st->mat[st->pmat[i]];
        switch (mt->model) {
            case 1: stress_rot1(st, i); f3dm1(st, i, mt->strinput); aviscosity(st, i); lieupd(st, i); break;
            case 3: stress_rot1(st, i); f3dm3(st, i, mt->strinput); aviscosity(st, i); lieupd(st, i); break;
            case 4: stress_rot1(st, i); f3dm25(st, i, mt->strinput); aviscosity(st, i); lieupd(st, i); break;
            case 9:
                stress_rot1(st, i); f3dm9(st, i, mt->strinput); sueos(st, i); aviscosity(st, i);
                hieupd(st, i); eqos(st, i); stress_up(st, i);
                break;
            case 10:
                stress_rot1(st, i); f3dm10(st, i, mt->strinput); sueos(st, i); aviscosity(st, i);
                hieupd(st, i); eqos(st, i); stress_up(st, i);
                break;
            default: set_err(st, ORC_ERR_MODEL); return st->error;
        }
    }
    calccritts(st);
    return st->error;
}

/* mcm_solution.f90:58-62 */
int orc_dt_update(orc_state *st) {
    st->dtold = st->dt;
    st->dt = st->tssf * st->critts;
    st->dt = fmin_(st->dt, 1.1 * st->dtold);
    st->ptime = st->ptime + st->dt;
    return st->error;
}

/* ------------------------------------------------------------ momentum */
/* mcm_update_ghost_stress.f90:28-41: +1 on the diagonal, -1 off it */
static void update_ghost_stress(orc_state *st) {
    for (int g = 0; g < st->ngp; g++) {
        int id = st->g_par[g];
        for (int c = 0; c < 3; c++)
            for (int r = 0; r < 3; r++) {
                double f = r == c ? 1.0 : -1.0;
                T9(st->g_sigma, g, r, c) = f * T9(st->sigma, id, r, c);
                T9(st->g_q, g, r, c) = f * T9(st->q, id, r, c);
            }
    }
}

static void nbr_moment_info(const orc_state *st, int i, int k, double *xj, double *mj, double *rhoj, double *hj,
                            const double **sgj, const double **qj) {
    if (k < st->nnbr[i]) {   /* mcm_get_j_moment_info.f90:31-58 */
        int j = st->nbr[(long)i * st->ncap + k];
        for (int d = 0; d < 3; d++) xj[d] = V3(st->x, j, d);
        *mj = st->mass[j]; *rhoj = st->rho[j]; *hj = st->h[j];
        *sgj = &st->sigma[9 * (long)j]; *qj = &st->q[9 * (long)j];
    } else {
        int g = st->gnbr[(long)i * st->gcap + (k - st->nnbr[i])];
        for (int d = 0; d < 3; d++) xj[d] = V3(st->g_x, g, d);
        *mj = st->g_mass[g]; *rhoj = st->g_rho[g]; *hj = st->g_h[g];
        *sgj = &st->g_sigma[9 * (long)g]; *qj = &st->g_q[9 * (long)g];
    }
}

/* mcm_oldacceleration.f90:36-90 */
static void oldacceleration(orc_state *st) {
#pragma omp parallel for schedule(dynamic, 256) if (st->np >= 16384)
    for (int i = 0; i < st->np; i++) {
        for (int d = 0; d < 3; d++) V3(st->a, i, d) = 0.0;
        if (st->contacttype > 0)   /* :49-56 */
            for (int d = 0; d < 3; d++) V3(st->a, i, d) = V3(st->a, i, d) + V3(st->rep, i, d);
        double rhoi = st->rho[i];
        const double *sgi = &st->sigma[9 * (long)i], *qi = &st->q[9 * (long)i];
        int nk = st->nnbr[i] + st->g_nnbr[i];
        for (int k = 0; k < nk; k++) {
            double xj[3], mj, rhoj, hj, dwdx[3], ds[3] = {0.0, 0.0, 0.0};
            const double *sgj, *qj;
            nbr_moment_info(st, i, k, xj, &mj, &rhoj, &hj, &sgj, &qj);
            double havg = 0.5 * (st->h[i] + hj);
            orc_kernel_gradw(&V3(st->x, i, 0), xj, havg, st->pi, dwdx);
            for (int n = 0; n < 3; n++)
                for (int m = 0; m < 3; m++)
                    ds[n] = ds[n] + ((sgj[m + 3 * n] - qj[m + 3 * n]) / (rhoj * rhoj) +
                                     (sgi[m + 3 * n] - qi[m + 3 * n]) / (rhoi * rhoi)) * dwdx[m];
            for (int l = 0; l < 3; l++) V3(st->a, i, l) = V3(st->a, i, l) + mj * ds[l];
        }
    }
}

/* mcm_corr_acceleration.f90:37-163: kernel centred at j, self term zeroed at
   k == nnbr, mass of the last neighbour visited, B^-1 contracted transposed */
static void corr_acceleration(orc_state *st) {
#pragma omp parallel for schedule(dynamic, 256) if (st->np >= 16384)
    for (int i = 0; i < st->np; i++) {
        for (int d = 0; d < 3; d++) V3(st->a, i, d) = 0.0;
        if (st->contacttype > 0)   /* :45-49 */
            for (int d = 0; d < 3; d++) V3(st->a, i, d) = V3(st->a, i, d) + V3(st->rep, i, d);
        const double *xi = &V3(st->x, i, 0);
        double rhoi = st->rho[i];
        const double *sgi = &st->sigma[9 * (long)i], *qi = &st->q[9 * (long)i];
        int nk = st->nnbr[i] + st->g_nnbr[i];
        double sum_w = 0.0, sum_gradw[3] = {0.0, 0.0, 0.0}, massj = 0.0;
        for (int k = 0; k < nk; k++) {
            double xj[3], mj, rhoj, hj, dwdx[3];
            const double *sgj, *qj;
            nbr_moment_info(st, i, k, xj, &mj, &rhoj, &hj, &sgj, &qj);
            double volj = mj / rhoj, havg = 0.5 * (st->h[i] + hj);
            double w = orc_kernel_w(xj, xi, havg, st->pi);
            orc_kernel_gradw(xj, xi, havg, st->pi, dwdx);
            sum_w = sum_w + volj * w;
            for (int l = 0; l < 3; l++) sum_gradw[l] = sum_gradw[l] + volj * dwdx[l];
            massj = mj;
        }
        double B[9] = {0}, invB[9];
        for (int k = 0; k < nk; k++) {
            double xj[3], mj, rhoj, hj, dwdx[3], cg[3];
            const double *sgj, *qj;
            nbr_moment_info(st, i, k, xj, &mj, &rhoj, &hj, &sgj, &qj);
            double volj = mj / rhoj, havg = 0.5 * (st->h[i] + hj);
            double w = orc_kernel_w(xj, xi, havg, st->pi);
            orc_kernel_gradw(xj, xi, havg, st->pi, dwdx);
            corr_grad(dwdx, w, sum_w, sum_gradw, cg);
            for (int m = 0; m < 3; m++)
                for (int n = 0; n < 3; n++) B[n + 3 * m] = B[n + 3 * m] + volj * (xj[n] - xi[n]) * cg[m];
        }
        orc_inverse3(B, invB);
        double dsig3[27] = {0};   /* deltasig(m,n,l) at m + 3n + 9l */
        for (int k = 0; k < nk; k++) {
            double xj[3], mj, rhoj, hj, dwdx[3], cg[3];
            const double *sgj, *qj;
            nbr_moment_info(st, i, k, xj, &mj, &rhoj, &hj, &sgj, &qj);
            double havg = 0.5 * (st->h[i] + hj);
            double w = orc_kernel_w(xj, xi, havg, st->pi);
            orc_kernel_gradw(xj, xi, havg, st->pi, dwdx);
            corr_grad(dwdx, w, sum_w, sum_gradw, cg);
            for (int n = 0; n < 3; n++)
                for (int m = 0; m < 3; m++) {
                    double dsig = (sgj[m + 3 * n] - qj[m + 3 * n]) / (rhoj * rhoj) +
                                  (sgi[m + 3 * n] - qi[m + 3 * n]) / (rhoi * rhoi);
                    if (k == st->nnbr[i] - 1) dsig = 0.0;
                    for (int l = 0; l < 3; l++) dsig3[m + 3 * n + 9 * l] = dsig3[m + 3 * n + 9 * l] + dsig * cg[l];
                }
        }
        for (int n = 0; n < 3; n++)
            for (int m = 0; m < 3; m++)
                for (int l = 0; l < 3; l++)
                    V3(st->a, i, n) = V3(st->a, i, n) + massj * dsig3[n + 3 * m + 9 * l] * invB[l + 3 * m];
    }
}

/* mcm_rep_contact.f90:20-98 with mcm_calc_fij (:100-161) and
   mcm_get_j_cont_info.f90:26-42: Monaghan's repulsive term over the real
   contact partners (other materials within 2h), B-spline normalisation at
   the partner's spacing, exponent n of particle i's material */
int orc_rep_cont(orc_state *st) {
#pragma omp parallel for schedule(dynamic, 256) if (st->np >= 16384)
    for (int i = 0; i < st->np; i++) {
        for (int d = 0; d < 3; d++) V3(st->rep, i, d) = 0.0;
        if (st->ncont[i] <= 0) continue;
        if (st->g_ncont[i] > 0) { set_err(st, ORC_ERR_UNDEFINED); continue; }
        const double ki = st->mat[st->pmat[i]].k_cont, n = st->mat[st->pmat[i]].n_cont;
        const double *xi = &V3(st->x, i, 0);
        for (int k = 0; k < st->ncont[i]; k++) {
            int j = st->cont[(long)i * st->ncap + k];
            const double kj = st->mat[st->pmat[j]].k_cont;
            const double *xj = &V3(st->x, j, 0);
            double massj = st->mass[j], hj = st->h[j];
            double havg = 0.5 * (st->h[i] + hj);
            double ratio = hj / havg, inv_w;
            if (ratio < 1.0) inv_w = 1.0 / (((0.75 * ratio - 1.5) * ratio) * ratio + 1.0);
            else { double t = 2.0 - ratio; inv_w = 1.0 / (t * (t * t)); }
            double dwdx[3];
            orc_kernel_gradw(xi, xj, havg, st->pi, dwdx);
            double rad = 0.0;
            for (int d = 0; d < 3; d++) { double xij = xi[d] - xj[d]; rad = rad + xij * xij; }
            rad = sqrt(rad);
            double z = rad / havg, fij;
            if (z < 1.0) fij = inv_w * (((0.75 * z - 1.5) * z) * z + 1.0);
            else if (z < 2.0) { double t = 2.0 - z; fij = 0.25 * inv_w * (t * (t * t)); }
            else fij = 0.0;
            double kij = ki + kj;
            for (int l = 0; l < 3; l++) V3(st->rep, i, l) = V3(st->rep, i, l) - massj * (kij * pow(fij, n)) * dwdx[l];
        }
    }
    return st->error;
}

/* mcm_momentum.f90:22-43 */
int orc_momentum(orc_state *st) {
    if (st->boundary) update_ghost_stress(st);
    if (st->disctype == 0) oldacceleration(st); else corr_acceleration(st);
    return st->error;
}

/* mcm_boundcond.f90:23-62 (no base acceleration) */
int orc_boundcond(orc_state *st) {
    static const int mask[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                   {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
    if (st->baseaccel)   /* :23-29 */
        for (int i = 0; i < st->np; i++)
            for (int d = 0; d < 3; d++) V3(st->a, i, d) = V3(st->a, i, d) + st->base_a[d];
    for (int i = 0; i < st->np; i++) {
        int bc = st->dispbc[i];
        if (bc < 1 || bc > 7) continue;
        for (int d = 0; d < 3; d++)
            if (mask[bc][d]) V3(st->a, i, d) = 0.0;
    }
    return st->error;
}

/* mcm_calctotale.f90:27-60 */
int orc_calctotale(orc_state *st) {
    for (int m = 0; m < st->nummat; m++) {
        st->mat_ie[m] = 0.0; st->mat_ke[m] = 0.0;
        for (int d = 0; d < 3; d++) st->mat_mom[3 * m + d] = 0.0;
    }
    for (int i = 0; i < st->np; i++) {
        int m = st->pmat[i];
        st->mat_ke[m] = st->mat_ke[m] + 0.5 * st->mass[i] * (st->vabs[i] * st->vabs[i]);
        for (int d = 0; d < 3; d++) st->mat_mom[3 * m + d] = st->mat_mom[3 * m + d] + st->mass[i] * fabs(V3(st->v, i, d));
    }
    for (int i = 0; i < st->np; i++) st->mat_ie[st->pmat[i]] = st->mat_ie[st->pmat[i]] + st->e[i];
    st->kinetice = 0.0; st->internale = 0.0;
    for (int m = 0; m < st->nummat; m++) {
        st->kinetice = st->kinetice + st->mat_ke[m];
        st->internale = st->internale + st->mat_ie[m];
    }
    st->totale = st->kinetice + st->internale;
    return st->error;
}

/* mcm_update_velocity.f90:29-87 (dynamic relaxation; XSPH when veloc_opt = 1) */
int orc_update_velocity(orc_state *st) {
    double dtn = 0.5 * (st->dt + st->dtold);
    for (int i = 0; i < st->np; i++) {
        double vabs = 0.0;
        for (int j = 0; j < 3; j++) {
            if (st->drelax)   /* :31-41 */
                V3(st->v, i, j) = st->drelax_scale * V3(st->v, i, j) + V3(st->a, i, j) * dtn;
            else
                V3(st->v, i, j) = V3(st->v, i, j) + V3(st->a, i, j) * dtn;
            vabs = vabs + V3(st->v, i, j) * V3(st->v, i, j);
        }
        st->vabs[i] = sqrt(vabs);
    }
    if (st->veloc_opt == 1) {
        /* XSPH, :61-87: every particle's smoothing from the new velocities,
           then all applied; the kernel at the partner's h, the volume over
           the mean density; |v| keeps the unsmoothed value */
        const double eps = 0.100;
        double *sv = xcalloc(3 * (size_t)st->np, 8);
        for (int i = 0; i < st->np; i++) {
            int nk = st->nnbr[i] + st->g_nnbr[i];
            for (int k = 0; k < nk; k++) {
                double xj[3], vj[3], mj, rhoj, hj;
                nbr_strain_info(st, i, k, xj, vj, &mj, &rhoj, &hj);
                double w = orc_kernel_w(&V3(st->x, i, 0), xj, hj, st->pi);
                double volj = mj / (0.5 * (rhoj + st->rho[i]));
                for (int d = 0; d < 3; d++) V3(sv, i, d) = V3(sv, i, d) + volj * (vj[d] - V3(st->v, i, d)) * w;
            }
        }
        for (int i = 0; i < st->np; i++)
            for (int d = 0; d < 3; d++) V3(st->v, i, d) = V3(st->v, i, d) + eps * V3(sv, i, d);
        free(sv);
    }
    return st->error;
}

/* mcm_check_sym_pen.f90:27-179 */
static void check_sym_pen(orc_state *st) {
    for (int d = 0; d < 3; d++) {
        double b0 = st->boundary_x[d][0], b1 = st->boundary_x[d][1];
        if (st->boundary_code[d][0] == 1 && st->coord_maxmin[d][0] < b0)
            for (int i = 0; i < st->np; i++)
                if (V3(st->x, i, d) < b0) { V3(st->x, i, d) = 2.0 * b0 - V3(st->x, i, d); V3(st->v, i, d) = -V3(st->v, i, d); }
        if (st->boundary_code[d][1] == 1 && st->coord_maxmin[d][1] > b1)
            for (int i = 0; i < st->np; i++)
                if (V3(st->x, i, d) > b1) { V3(st->x, i, d) = 2.0 * b1 - V3(st->x, i, d); V3(st->v, i, d) = -V3(st->v, i, d); }
        if (st->boundary_code[d][0] == 2) {
            double span = b1 - b0;
            if (st->coord_maxmin[d][0] < b0)
                for (int i = 0; i < st->np; i++)
                    if (V3(st->x, i, d) < b0) V3(st->x, i, d) = V3(st->x, i, d) + span;
            if (st->coord_maxmin[d][1] > b1)
                for (int i = 0; i < st->np; i++)
                    if (V3(st->x, i, d) > b1) V3(st->x, i, d) = V3(st->x, i, d) - span;
        }
    }
}

/* mcm_move.f90:25-55 */
int orc_move(orc_state *st) {
    for (int d = 0; d < 3; d++) { st->coord_maxmin[d][0] = 1.0e20; st->coord_maxmin[d][1] = -1.0e20; }
    for (int i = 0; i < st->np; i++)
        for (int j = 0; j < 3; j++) {
            V3(st->x, i, j) = V3(st->x, i, j) + V3(st->v, i, j) * st->dt;
            if (V3(st->x, i, j) < st->coord_maxmin[j][0]) st->coord_maxmin[j][0] = V3(st->x, i, j);
            if (V3(st->x, i, j) > st->coord_maxmin[j][1]) st->coord_maxmin[j][1] = V3(st->x, i, j);
        }
    if (st->boundary) check_sym_pen(st);
    return st->error;
}

/* ------------------------------------------------------ initialisation */
/* mcm_eos{1,4,13,28,41}p.f90 (initial pressure) */
static void eos_p(orc_state *st, int i) {
    const orc_material *mt = &st->mat[st->pmat[i]];
    const double *ei = mt->eosinput;
    double rho = st->rho[i], rho0 = st->rho0[i];
    switch (mt->eos) {
        case 1: {
            double mu = rho / rho0 - 1.0, mu_t = fmax_(mu, 0.0), e1 = 0.0;
            st->p[i] = ei[0] + (ei[1] + ei[3] * (mu * mu)) * mu + ei[2] * (mu_t * mu_t) +
                       (ei[4] + ei[5] * mu + ei[6] * (mu_t * mu_t)) * e1;
            break;
        }
        case 4: {
            double c1 = ei[0], s1 = ei[1], s2 = ei[2], s3 = ei[3], gamma = ei[4], a1 = ei[5];
            double e1 = st->e[i] * rho0 / st->mass[i];
            double mu = rho / rho0 - 1.0;
            if (mu > 0.0) {
                double termone = mu * (1.0 - (gamma / 2.0));
                double termtwo = (a1 / 2.0) * mu * mu;
                double eosnum = mu * rho0 * (c1 * c1) * (1.0 + termone - termtwo);
                double mp1 = mu + 1.0;
                double mu21 = (mu * mu) / mp1;
                double mu32 = (mu * mu * mu) / (mp1 * mp1);
                double eosden = 1.0 - (s1 - 1.0) * mu - s2 * mu21 - s3 * mu32;
                st->p[i] = eosnum / (eosden * eosden) + (gamma + a1 * mu) * e1;
            } else {
                st->p[i] = rho0 * (c1 * c1) * mu + (gamma + a1 * mu) * e1;
            }
            break;
        }
        case 13: st->p[i] = rho * (ei[0] - 1.0) * (st->e[i] / st->mass[i]); break;
        case 28: st->p[i] = ei[0] * (pow(rho / rho0, ei[1]) - 1.0); break;
        case 41: {
            double cc = ei[0], sl = ei[1], gamma = ei[4];
            double e1 = st->e[i] / st->mass[i];
            double ao = rho0 * (cc * cc);
            double mu = rho / rho0 - 1.0;
            double slm = sl - 1.0;
            double bo = ao * (1.0 + 2.0 * (sl - 1.0));
            double co = ao * (2.0 * (sl - 1.0) + 3.0 * (slm * slm));
            gamma = gamma * rho0 / rho;
            double f = gamma * mu, ph;
            if (mu > 0.0) ph = ao * mu + bo * (mu * mu) + co * (mu * (mu * mu));
            else ph = ao * mu;
            double pn = ph * (1.0 - 0.5 * f) + rho * gamma * e1;
            if (pn < st->pcut[i]) {
                pn = st->pcut[i];
                st->rho[i] = st->rho[i] / (1.0 - st->tracerod[i] * st->dt);
            }
            st->p[i] = pn;
            break;
        }
        default: break;
    }
}

/* mcm_eoscalc.f90:27-113 */
static void eoscalc(orc_state *st, int i) {
    const orc_material *mt = &st->mat[st->pmat[i]];
    const double *ei = mt->eosinput;
    if (st->init_rhoe != 1) {
        switch (mt->eos) {
            case 1: st->rho[i] = st->rho0[i] / ei[8]; st->e[i] = ei[7] * st->mass[i] / st->rho[i]; break;
            case 4: st->rho[i] = st->rho0[i] / ei[7]; st->e[i] = ei[6] * st->mass[i] / st->rho[i]; break;
            case 13: st->rho[i] = st->rho0[i] / ei[3]; st->e[i] = ei[2] * st->mass[i] / st->rho[i]; break;
            case 28: st->rho[i] = st->rho0[i] / ei[3]; st->e[i] = ei[2] * st->mass[i]; break;
            case 41: st->rho[i] = st->rho0[i] / ei[5]; st->e[i] = ei[2] * ei[3] * st->mass[i] / st->rho[i]; break;
            default: set_err(st, ORC_ERR_MODEL); break;
        }
    }
    eos_p(st, i);
    sueos(st, i);
}

/* mcm_initial.f90:33-101 without the output calls */
int orc_initial(orc_state *st) {
    static const int mask[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                   {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
    int n = st->np;
    for (int i = 0; i < n; i++) {   /* mcm_init_boundcond.f90:25-54 */
        int bc = st->dispbc[i];
        if (bc >= 1 && bc <= 7)
            for (int d = 0; d < 3; d++)
                if (mask[bc][d]) V3(st->v, i, d) = 0.0;
    }
    if (st->massopt == 0) {          /* mcm_calcmass.f90:122-137 */
        int *cnt = xcalloc(st->nummat, sizeof(int));
        for (int i = 0; i < n; i++) cnt[st->pmat[i]]++;
        for (int i = 0; i < n; i++) st->mass[i] = st->mat[st->pmat[i]].mass / (double)cnt[st->pmat[i]];
        free(cnt);
    }
    for (int i = 0; i < n; i++) {    /* mcm_initvbls.f90:23-91 */
        for (int d = 0; d < 3; d++) { V3(st->xzero, i, d) = V3(st->x, i, d); V3(st->a, i, d) = 0.0; }
        st->efps[i] = 0.0; st->fail[i] = 1.0;
        int model = st->mat[st->pmat[i]].model;
        if (model == 1 || model == 3 || model == 4) st->pcut[i] = -9.90E+20;
        else if (model == 9) st->pcut[i] = st->mat[st->pmat[i]].strinput[0];
        else if (model == 10) st->pcut[i] = st->mat[st->pmat[i]].strinput[3];
        double vabs = 0.0;
        for (int d = 0; d < 3; d++) vabs = vabs + V3(st->v, i, d) * V3(st->v, i, d);
        st->vabs[i] = sqrt(vabs);
    }
    for (int d = 0; d < 3; d++) { st->coord_maxmin[d][0] = 1.0e20; st->coord_maxmin[d][1] = -1.0e20; }
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            if (V3(st->x, i, d) < st->coord_maxmin[d][0]) st->coord_maxmin[d][0] = V3(st->x, i, d);
            if (V3(st->x, i, d) > st->coord_maxmin[d][1]) st->coord_maxmin[d][1] = V3(st->x, i, d);
        }
    for (int m = 0; m < st->nummat; m++) {   /* mcm_calcg.f90:31-59 */
        orc_material *mt = &st->mat[m];
        switch (mt->model) {
            case 1: case 3: mt->g = mt->strinput[0] / (2.0 * (1.0 + mt->strinput[5])); break;
            case 4: mt->g = mt->strinput[5]; break;
            case 9: mt->g = 0.0; break;
            case 10: mt->g = mt->strinput[0]; break;
        }
    }
    for (int i = 0; i < n; i++) {    /* mcm_calcrhopc.f90:31-95 */
        const orc_material *mt = &st->mat[st->pmat[i]];
        st->rho0[i] = mt->rho;
        switch (mt->model) {
            case 1: case 3:
                if (st->init_rhoe != 1) { st->rho[i] = st->rho0[i]; st->e[i] = 0.0; }
                st->p[i] = 0.0;
                st->c[i] = sqrt(mt->strinput[1] / st->rho[i]);
                break;
            case 4:   /* :57-67, the sound speed from lambda + 2G of mcm_sets4 */
                if (st->init_rhoe != 1) { st->rho[i] = st->rho0[i]; st->e[i] = 0.0; }
                st->p[i] = 0.0;
                st->c[i] = sqrt(mt->strinput[9] / st->rho[i]);
                break;
            case 9:
                eoscalc(st, i);
                if (st->p[i] < mt->strinput[0]) st->p[i] = mt->strinput[0];
                break;
            case 10: eoscalc(st, i); break;
            default: set_err(st, ORC_ERR_MODEL); break;
        }
    }
    for (int i = 0; i < n; i++) {    /* mcm_initsigma.f90 / mcm_initrod.f90 */
        for (int k = 0; k < 9; k++) {
            st->sigma[9 * (long)i + k] = 0.0; st->s[9 * (long)i + k] = 0.0; st->alfa[9 * (long)i + k] = 0.0;
            st->rod[9 * (long)i + k] = 0.0; st->spin[9 * (long)i + k] = 0.0;
        }
        for (int d = 0; d < 3; d++) T9(st->sigma, i, d, d) = -st->p[i];
    }
    st->hmax = 0.0;                  /* mcm_init_h.f90:22-48 */
    if (st->init_h_opt == 0) {
        for (int i = 0; i < n; i++) st->h[i] = st->mat[st->pmat[i]].h;
        for (int m = 0; m < st->nummat; m++) st->hmax = fmax_(st->hmax, st->mat[st->pmat[m]].h);
    } else {
        for (int i = 0; i < n; i++) st->hmax = fmax_(st->hmax, st->h[i]);
    }
    for (int i = 0; i < n; i++) { st->h0[i] = st->h[i]; st->hold[i] = st->h[i]; }
    st->timestep = 1;
    for (int i = 0; i < n; i++) {
        st->tracerod[i] = 0.0;
        for (int l = 0; l < 3; l++) st->tracerod[i] = st->tracerod[i] + T9(st->rod, i, l, l);
        aviscosity(st, i);
    }
    setuplist(st);                   /* mcm_init_neighb.f90 */
    if (st->boundary) { init_boundary(st); ghost_setup(st); }
    ll_neighbours(st, st->disctype == 1 ? 1 : 0);
    calccritts(st);
    if (st->itss != 0.0) st->dt = fmin_(st->itss, st->critts * st->tssf);
    else st->dt = st->critts * st->tssf;
    orc_calctotale(st);
    for (int i = 0; i < n; i++) {    /* mcm_initold.f90 */
        st->rhoold[i] = st->rho[i];
        for (int k = 0; k < 9; k++) st->qold[9 * (long)i + k] = st->q[9 * (long)i + k];
    }
    return st->error;
}

int orc_step(orc_state *st) {
    orc_neighbours(st);
    orc_strain(st);
    if (st->h_opt > 0) orc_update_h(st);
    orc_constitutive(st);
    orc_dt_update(st);
    if (st->contacttype > 0) orc_rep_cont(st);   /* mcm_solution.f90:64 */
    orc_momentum(st);
    orc_boundcond(st);
    orc_calctotale(st);
    orc_update_velocity(st);
    orc_move(st);
    st->timestep++;
    return st->error;
}

int orc_run(orc_state *st, int nsteps) {
    for (int k = 0; k < nsteps && st->error == 0; k++) orc_step(st);
    return st->error;
}

/* ----------------------------------------------------------------- dump */
/* Same layout as ref_probe.f90 probe_dump (1-based ids). */
/* particle row of the dump layout (oracle/dumpio.py FIELDS) */
static void state_row(const orc_state *st, int i, double *row) {
    for (int d = 0; d < 3; d++) { row[d] = V3(st->x, i, d); row[3 + d] = V3(st->v, i, d); row[6 + d] = V3(st->a, i, d); }
    row[9] = st->rho[i]; row[10] = st->rhoold[i]; row[11] = st->h[i]; row[12] = st->hold[i];
    row[13] = st->p[i]; row[14] = st->e[i]; row[15] = st->etry[i]; row[16] = st->einc[i];
    row[17] = st->c[i]; row[18] = st->vabs[i]; row[19] = st->tracerod[i]; row[20] = st->pcut[i];
    row[21] = st->p_cut[i]; row[22] = st->efps[i]; row[23] = st->fail[i]; row[24] = st->mindist[i];
    row[25] = st->critts_p[i]; row[26] = st->mass[i]; row[27] = st->rho0[i];
    for (int k = 0; k < 9; k++) {
        row[28 + k] = st->sigma[9 * (long)i + k]; row[37 + k] = st->s[9 * (long)i + k];
        row[46 + k] = st->alfa[9 * (long)i + k]; row[55 + k] = st->q[9 * (long)i + k];
        row[64 + k] = st->rod[9 * (long)i + k]; row[73 + k] = st->spin[9 * (long)i + k];
    }
    row[82] = st->h0[i];
    for (int d = 0; d < 3; d++) row[83 + d] = V3(st->xzero, i, d);
}

int orc_dump(orc_state *st, const char *path, int step, int stage) {
    FILE *f = fopen(path, "wb");
    if (!f) return ORC_ERR_IO;
    int32_t hdr[6] = {0x4D434D44, step, stage, st->np, st->ngp, st->nummat};
    fwrite(hdr, 4, 6, f);
    double sc[8] = {st->dt, st->dtold, st->ptime, st->critts, st->hmax, st->kinetice, st->internale, st->totale};
    fwrite(sc, 8, 8, f);
    double g[12];
    for (int d = 0; d < 3; d++) { g[2 * d] = st->coord_maxmin[d][0]; g[2 * d + 1] = st->coord_maxmin[d][1]; }
    for (int d = 0; d < 3; d++) { g[6 + d] = st->gridmin[d]; g[9 + d] = st->gridsize[d]; }
    fwrite(g, 8, 12, f);
    fwrite(st->gridlim, 4, 3, f);
    for (int m = 0; m < st->nummat; m++) {
        double mm[5] = {st->mat_ke[m], st->mat_ie[m], st->mat_mom[3 * m], st->mat_mom[3 * m + 1], st->mat_mom[3 * m + 2]};
        fwrite(mm, 8, 5, f);
    }
    double row[86];
    for (int i = 0; i < st->np; i++) {
        state_row(st, i, row);
        fwrite(row, 8, 86, f);
    }
    fwrite(st->nnbr, 4, st->np, f);
    fwrite(st->g_nnbr, 4, st->np, f);
    for (int i = 0; i < st->np; i++) {
        int32_t v = st->llpointer[i];
        fwrite(&v, 4, 1, f);
    }
    for (int i = 0; i < st->np; i++)
        for (int k = 0; k < st->nnbr[i]; k++) {
            int32_t v = st->nbr[(long)i * st->ncap + k] + 1;
            fwrite(&v, 4, 1, f);
        }
    if (st->boundary) {
        for (int i = 0; i < st->np; i++)
            for (int k = 0; k < st->g_nnbr[i]; k++) {
                int32_t v = st->gnbr[(long)i * st->gcap + k] + 1;
                fwrite(&v, 4, 1, f);
            }
        for (int g = 0; g < st->ngp; g++) { int32_t v = st->g_par[g] + 1; fwrite(&v, 4, 1, f); }
        for (int g = 0; g < st->ngp; g++) { int32_t v = st->g_mat[g] + 1; fwrite(&v, 4, 1, f); }
        for (int g = 0; g < st->ngp; g++) {
            double r8[8] = {V3(st->g_x, g, 0), V3(st->g_x, g, 1), V3(st->g_x, g, 2), V3(st->g_v, g, 0),
                            V3(st->g_v, g, 1), V3(st->g_v, g, 2), st->g_mass[g], st->g_h[g]};
            fwrite(r8, 8, 8, f);
        }
        fwrite(st->g_rho, 8, st->ngp, f);
        fwrite(st->g_hold, 8, st->ngp, f);
    }
    int cap = 0;
    for (int m = 0; m < st->nummat; m++) cap |= st->mat[m].model == 4;
    if (cap) {   /* cap-model section: 'CAP4', capa, ptype */
        int32_t tag = 0x43415034;
        fwrite(&tag, 4, 1, f);
        fwrite(st->capa, 8, st->np, f);
        for (int i = 0; i < st->np; i++) { int32_t v = st->ptype[i]; fwrite(&v, 4, 1, f); }
    }
    fclose(f);
    return ORC_OK;
}

/* In-memory views for runs too large to dump: rows[np * 86] in the dump's
   row layout; lists as ids[np * cap] row-major, 0-based, ghosts as
   np + ghost id (real part first, then ghosts, in list order), with the
   real (incl. self) and ghost counts; scalars dt, dtold, ptime, critts,
   hmax, kinetice, internale, totale, ngp. */
void orc_get_rows(const orc_state *st, double *rows) {
#pragma omp parallel for schedule(static) if (st->np >= 16384)
    for (int i = 0; i < st->np; i++) state_row(st, i, rows + 86 * (long)i);
}

int orc_get_lists(const orc_state *st, int32_t *counts, int32_t *gcounts, int32_t *ids, int64_t cap) {
    for (int i = 0; i < st->np; i++)
        if (st->nnbr[i] + st->g_nnbr[i] > cap) return ORC_ERR_CAPACITY;
#pragma omp parallel for schedule(static) if (st->np >= 16384)
    for (int i = 0; i < st->np; i++) {
        counts[i] = st->nnbr[i];
        gcounts[i] = st->g_nnbr[i];
        int32_t *row = ids + (long)i * cap;
        for (int k = 0; k < st->nnbr[i]; k++) row[k] = st->nbr[(long)i * st->ncap + k];
        for (int k = 0; k < st->g_nnbr[i]; k++) row[st->nnbr[i] + k] = st->np + st->gnbr[(long)i * st->gcap + k];
    }
    return ORC_OK;
}

void orc_get_scalars(const orc_state *st, double *out) {
    const double v[9] = {st->dt, st->dtold, st->ptime, st->critts, st->hmax, st->kinetice, st->internale,
                         st->totale, (double)st->ngp};
    for (int k = 0; k < 9; k++) out[k] = v[k];
}

void orc_get_capa(const orc_state *st, double *capa, int32_t *ptype) {
    for (int i = 0; i < st->np; i++) { capa[i] = st->capa[i]; ptype[i] = st->ptype[i]; }
}

void orc_get_x(const orc_state *st, double *out) {
    for (long k = 0; k < 3L * st->np; k++) out[k] = st->x[k];
}
