/*
 * elec_oracle.c -- CPU restatement of the MB-pol polarizable electrostatics for one force
 * evaluation, with its per-atom energy / stress / force-matrix bookkeeping.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Site layout: waters first (O, H1, H2 of water w at 3w, 3w+1, 3w+2, 0-based), then the
 * M sites (M of water w at 3*nw + w), the layout ttm2list.f produces for a FIELD whose
 * water is the second-to-last molecule type and the M site the last.  Single rank, Verlet
 * half lists (so the eps / eps2 factors of forces.f:306-308 are epsq = 1), no second
 * exclusion list (lexatm2 is only filled by exclude_link.f for link-cell lists, so
 * ewald4p and rfield.f:531-585 do nothing here).
 *
 * Restated (reference file:line):
 *   M-site positions               dlpoly_cmd.f:2041-2057
 *   DMS charges and derivatives    qdetermine.f90:110-259 (table module.f90:765-812)
 *   erfc                            erfcc.f:1-11 (default-REAL literals: float values)
 *   B-function table               erfcgenp.f:29-49, grid size parset.f:210-217
 *   incomplete gamma               gammq.f:1-95 (EPS, FPMIN default REAL: float values)
 *   real-space fields              rfield.f:62-323 (list pairs), 327-525 (excluded pairs)
 *   reciprocal space               ewald1p.f:96-1030
 *   SOR update                     iter.f:45-101
 *   solve loop                     forces.f:268-404
 *   polarization energy            coulomb_polar.f:58-103
 *   real-space energy / forces     ewald2p.f:86-640
 *   excluded-pair corrections      ewald3p.f:76-391
 *   ESP sum, charge-derivative     forces.f:874-892, qdforce.f90:13-569
 *   heat bookkeeping               update_energy_*, update_stress_ew1p, update_forces
 *                                  (heatcurrent.f90:155-267,529-533), M-site
 *                                  redistribution dlpoly_cmd.f:2577-2606 and
 *                                  heatcurrent.f90:302-555
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define R4PIE0 138935.4835          /* dl_params.inc */
#define SQRPI 1.7724538509055159    /* dl_params.inc sqrpi */
#define PI_ 3.141592653589793
#define BOHR 0.529177249            /* module.f90:40 */

/* ------------------------------------------------------------------ specials ---- */

/* erfcc.f: Numerical Recipes erfc.  The literals are default REAL, i.e. single precision
 * values promoted to double. */
double ora_erfcc(double x)
{
#ifdef ORA_EXACT_ERFC
    /* the second build of this oracle (liboracle_exact.so): the C library's erfc in place of
     * the reference's approximation, to separate what that approximation does from the rest */
    return erfc(x);
#endif
    const double z = fabs(x);
    const double t = 1.0 / (1.0 + 0.5 * z);
    double r = t * exp(-z * z - (double)1.26551223f +
                       t * ((double)1.00002368f +
                            t * ((double).37409196f +
                                 t * ((double).09678418f +
                                      t * ((double)-.18628806f +
                                           t * ((double).27886807f +
                                                t * ((double)-1.13520398f +
                                                     t * ((double)1.48851587f +
                                                          t * ((double)-0.82215223f + t * (double).17087277f)))))))));
    if (x < 0.0) r = 2.0 - r;
    return r;
}

/* gammq.f gammln (Lanczos, NR) */
double ora_gammln(double xx)
{
    static const double cof[6] = {76.18009172947146, -86.50532032941677, 24.01409824083091,
                                  -1.231739572450155, .1208650973866179e-2, -.5395239384953e-5};
    const double stp = 2.5066282746310005;
    double x = xx, y = x;
    double tmp = x + 5.5;
    tmp = (x + 0.5) * log(tmp) - tmp;
    double ser = 1.000000000190015;
    for (int j = 0; j < 6; ++j) {
        y = y + 1.0;
        ser = ser + cof[j] / y;
    }
    return tmp + log(stp * ser / x);
}

/* gammq.f gser / gcf / gammq; EPS and FPMIN are REAL parameters (float values) */
static double gser(double a, double x)
{
    const double EPS = (double)3.e-7f;
    const double gln = ora_gammln(a);
    if (x <= 0.0) return 0.0;
    double ap = a, sum = 1.0 / a, del = sum;
    for (int n = 1; n <= 800; ++n) {
        ap = ap + 1.0;
        del = del * x / ap;
        sum = sum + del;
        if (fabs(del) < fabs(sum) * EPS) break;
    }
    return sum * exp(-x + a * log(x) - gln);
}

static double gcf(double a, double x)
{
    const double EPS = (double)3.e-7f, FPMIN = (double)1.e-30f;
    const double gln = ora_gammln(a);
    double b = x + 1.0 - a;
    double c = 1.0 / FPMIN;
    double d = 1.0 / b;
    double h = d;
    for (int i = 1; i <= 100; ++i) {
        const double an = -i * (i - a);
        b = b + 2.0;
        d = an * d + b;
        if (fabs(d) < FPMIN) d = FPMIN;
        c = b + an / c;
        if (fabs(c) < FPMIN) c = FPMIN;
        d = 1.0 / d;
        const double del = d * c;
        h = h * del;
        if (fabs(del - 1.0) < EPS) break;
    }
    return exp(-x + a * log(x) - gln) * h;
}

double ora_gammq(double a, double x)
{
    if (x < a + 1.0) return 1.0 - gser(a, x);
    return gcf(a, x);
}

/* erfcgenp.f: Smith B-functions (modified to give -erf) on r_i = i*drewd, i = 1..mxegrd;
 * row 0 is unused and left zero.  Grid: parset.f:210-217. */
int ora_erfc_grid(double rcut, double *drewd)
{
    int mxegrd = (int)(rcut / 0.0001 + 0.5) + 4;
    if (mxegrd < 1000) mxegrd = 1000;
    *drewd = rcut / (double)(mxegrd - 4);
    return mxegrd;
}

void ora_erfc_table(double alpha, double rcut, double *t /* (mxegrd+1)*4 */)
{
    double drewd;
    const int mxegrd = ora_erfc_grid(rcut, &drewd);
    memset(t, 0, sizeof(double) * 4);
    for (int i = 1; i <= mxegrd; ++i) {
        const double rrr = (double)i * drewd;
        const double rsqi = 1.0 / rrr / rrr;
        const double a = alpha * rrr;
        double *e = t + 4 * i;
        if (alpha < 1.0e-6) { /* no Ewald split (a cluster): the -erf kernels vanish, as ewald3p.f sets them */
            e[0] = e[1] = e[2] = e[3] = 0.0;
            continue;
        }
        e[0] = (ora_erfcc(a) - 1.0) / rrr;
        const double exp2a = exp(-a * a);
        for (int n = 1; n <= 3; ++n) {
            const double fn = (double)n;
            e[n] = rsqi * ((2.0 * fn - 1.0) * e[n - 1] + pow(2.0 * alpha * alpha, fn) * exp2a / alpha / SQRPI);
        }
    }
}

/* 3-point interpolation of the table (rfield.f:102-120, ewald2p.f:160-178) */
static void table_interp(const double *t, double drewd, double rr, double *bbb)
{
    const int ll = (int)(rr / drewd);
    const double ppp = rr / drewd - (double)ll;
    for (int n = 0; n < 4; ++n) {
        const double vk = t[4 * ll + n], vk1 = t[4 * (ll + 1) + n], vk2 = t[4 * (ll + 2) + n];
        const double t1 = vk + (vk1 - vk) * ppp;
        const double t2 = vk1 + (vk2 - vk1) * (ppp - 1.0);
        bbb[n] = t1 + (t2 - t1) * ppp * 0.5;
    }
}

/* direct B functions of excluded pairs (rfield.f:365-374, ewald3p.f:100-110) */
static void bfun_direct(double alpha, double rr, double rsqi, double *bbb)
{
    if (alpha < 1.0e-6) { /* ewald3p.f: "if ewald parameter is too small then turn it off" */
        bbb[0] = bbb[1] = bbb[2] = bbb[3] = 0.0;
        return;
    }
    const double a = alpha * rr;
    bbb[0] = (ora_erfcc(a) - 1.0) / rr;
    const double exp2a = exp(-a * a);
    for (int m = 1; m <= 3; ++m) {
        const double fm = (double)m;
        bbb[m] = rsqi * ((2.0 * fm - 1.0) * bbb[m - 1] + pow(2.0 * pow(alpha, 2.0), fm) * exp2a / alpha / SQRPI);
    }
}

/* ----------------------------------------------------------------- DMS table ---- */

static int g_dms_loaded = 0;
static int g_dms_idx[84][3];
static double g_dms_coef[84];

int ora_dms_load(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    for (int j = 0; j < 84; ++j)
        if (fscanf(f, "%d %d %d %lf", &g_dms_idx[j][0], &g_dms_idx[j][1], &g_dms_idx[j][2], &g_dms_coef[j]) != 4) {
            fclose(f);
            return -2;
        }
    fclose(f);
    g_dms_loaded = 1;
    return 0;
}

/* DMS parameters (module.f90:818-826) and the reference geometry of the expansion
 * (ps_type_pes reoh, thetae; module.f90:480-481, cose = cos(thetae) :596) */
#define DMS_B1 1.0
#define DMS_A (0.2999 * BOHR)
#define DMS_B (-0.6932)
#define DMS_C0 1.0099
#define DMS_C1 (-0.1801)
#define DMS_C2 0.0892
#define PS_REOH 0.958649
#define PS_THETAE (104.3475 * PI_ / 180.0)

typedef struct {
    double dp1dr1, dp1dr2, dp2dr1, dp2dr2, dp1dcos, dp2dcos;
} dms_grad_t;

/* qdetermine.f90:118-259 for one water: p1, p2 and their derivatives; the new charge
 * assignment (lttmfold = .false., lttm3 = .false. for MB-pol, simdef_cmd.f vmbpol). */
static void dms_water(int imcon, const double *cell, const double *o, const double *h1, const double *h2,
                      double gamma, double *qh1, double *qh2, dms_grad_t *G, double *pout)
{
    double r1v[3] = {h1[0] - o[0], h1[1] - o[1], h1[2] - o[2]};
    double r2v[3] = {h2[0] - o[0], h2[1] - o[1], h2[2] - o[2]};
    ora_image_vectors(imcon, cell, 1, r1v);
    ora_image_vectors(imcon, cell, 1, r2v);
    const double r1 = sqrt(r1v[0] * r1v[0] + r1v[1] * r1v[1] + r1v[2] * r1v[2]);
    const double r2 = sqrt(r2v[0] * r2v[0] + r2v[1] * r2v[1] + r2v[2] * r2v[2]);
    const double cost = (r1v[0] * r2v[0] + r1v[1] * r2v[1] + r1v[2] * r2v[2]) / (r1 * r2);
    const double reoh = PS_REOH, cose = cos(PS_THETAE);
    const double x1 = (r1 - reoh) / reoh, x2 = (r2 - reoh) / reoh, x3 = cost - cose;
    double fm[11][3];
    memset(fm, 0, sizeof fm);
    fm[1][0] = fm[1][1] = fm[1][2] = 1.0;
    int isump = 0;
    for (int j = 0; j < 84; ++j)
        for (int k = 0; k < 3; ++k)
            if (g_dms_idx[j][k] > isump) isump = g_dms_idx[j][k];
    for (int j = 2; j <= isump; ++j) {
        fm[j][0] = fm[j - 1][0] * x1;
        fm[j][1] = fm[j - 1][1] * x2;
        fm[j][2] = fm[j - 1][2] * x3;
    }
    double p1 = 0.0, p2 = 0.0;
    dms_grad_t g = {0, 0, 0, 0, 0, 0};
    for (int j = 1; j < 84; ++j) {
        const int a = g_dms_idx[j][0], b = g_dms_idx[j][1], c = g_dms_idx[j][2];
        const double cf = g_dms_coef[j];
        p1 = p1 + cf * fm[a][0] * fm[b][1] * fm[c][2];
        p2 = p2 + cf * fm[b][0] * fm[a][1] * fm[c][2];
        g.dp1dr1 = g.dp1dr1 + cf * (a - 1) * fm[a - 1][0] * fm[b][1] * fm[c][2] / reoh;
        g.dp1dr2 = g.dp1dr2 + cf * (b - 1) * fm[a][0] * fm[b - 1][1] * fm[c][2] / reoh;
        g.dp2dr1 = g.dp2dr1 + cf * (b - 1) * fm[b - 1][0] * fm[a][1] * fm[c][2] / reoh;
        g.dp2dr2 = g.dp2dr2 + cf * (a - 1) * fm[b][0] * fm[a - 1][1] * fm[c][2] / reoh;
        g.dp1dcos = g.dp1dcos + cf * (c - 1) * fm[a][0] * fm[b][1] * fm[c - 1][2];
        g.dp2dcos = g.dp2dcos + cf * (c - 1) * fm[b][0] * fm[a][1] * fm[c - 1][2];
    }
    const double pl1 = cost, pl2 = 0.5 * (3.0 * pl1 * pl1 - 1.0);
    const double damp = exp(-DMS_B1 * ((r1 - reoh) * (r1 - reoh) + (r2 - reoh) * (r2 - reoh)));
    const double ddampdr1 = -2.0 * DMS_B1 * (r1 - reoh) * damp;
    const double ddampdr2 = -2.0 * DMS_B1 * (r2 - reoh) * damp;
    const double leg = DMS_C0 + pl1 * DMS_C1 + pl2 * DMS_C2;
    const double pc0 = DMS_A * (pow(r1, DMS_B) + pow(r2, DMS_B)) * leg;
    const double dpc0dr1 = DMS_A * (DMS_B * pow(r1, DMS_B - 1)) * leg;
    const double dpc0dr2 = DMS_A * (DMS_B * pow(r2, DMS_B - 1)) * leg;
    const double dpc0dcos = DMS_A * (pow(r1, DMS_B) + pow(r2, DMS_B)) * (DMS_C1 + 3.0 * DMS_C2 * pl1);
    g.dp1dr1 = g.dp1dr1 * damp + p1 * ddampdr1 + dpc0dr1;
    g.dp1dr2 = g.dp1dr2 * damp + p1 * ddampdr2 + dpc0dr2;
    g.dp2dr1 = g.dp2dr1 * damp + p2 * ddampdr1 + dpc0dr1;
    g.dp2dr2 = g.dp2dr2 * damp + p2 * ddampdr2 + dpc0dr2;
    g.dp1dcos = g.dp1dcos * damp + dpc0dcos;
    g.dp2dcos = g.dp2dcos * damp + dpc0dcos;
    p1 = g_dms_coef[0] + p1 * damp + pc0;
    p2 = g_dms_coef[0] + p2 * damp + pc0;
    const double factor = gamma / 2.0 / (1.0 - gamma);
    *qh1 = p1 + factor * (p1 + p2);
    *qh2 = p2 + factor * (p1 + p2);
    if (G) *G = g;
    if (pout) {
        pout[0] = p1;
        pout[1] = p2;
    }
}

int ora_dms_charges(int imcon, const double *cell, const double *w9, double gamma, double *q3)
{
    if (!g_dms_loaded) return -1;
    double qh1, qh2;
    dms_water(imcon, cell, w9, w9 + 3, w9 + 6, gamma, &qh1, &qh2, NULL, NULL);
    q3[0] = qh1;
    q3[1] = qh2;
    q3[2] = -(qh1 + qh2);
    return 0;
}

int ora_dms_p(int imcon, const double *cell, const double *w9, double *p8)
{
    if (!g_dms_loaded) return -1;
    double qh1, qh2;
    dms_grad_t g;
    dms_water(imcon, cell, w9, w9 + 3, w9 + 6, 0.0, &qh1, &qh2, &g, p8);
    p8[2] = g.dp1dr1;
    p8[3] = g.dp1dr2;
    p8[4] = g.dp1dcos;
    p8[5] = g.dp2dr1;
    p8[6] = g.dp2dr2;
    p8[7] = g.dp2dcos;
    return 0;
}

/* ------------------------------------------------------------------ context ---- */

typedef struct {
    ora_elec_args *A;
    int nw, n, nreal;
    double *chge, *polr, *polr2, *dip;
    double *x, *y, *z;
    /* fields */
    double *efk, *emu, *efdc, *efdd;
    /* ESP parts */
    double *vk, *vr, *vs, *vc, *vdck, *vdcr;
    /* reciprocal space */
    double rc[9], volm, rcpct2, rvolm, ralph, engsic;
    double *elc1, *els1, *emc, *ems, *enc, *ens, *elc0, *els0, *ckr, *skr, *fk;
    /* real space */
    double *tab, drewd;
    long npair;
    int *pi, *pj;
    dms_grad_t *dg;
} ctx_t;

static void img(const ctx_t *C, double *d)
{
    ora_image_vectors(C->A->imcon, C->A->cell, 1, d);
}

/* force-matrix entry F at (row a, column b), i.e. update_forces(a, b, F): dense copy
 * (optional) and the per-atom contraction W_r += w F (x) (x_r - x_b) over real columns
 * b with the M-site rows moved to O, H1, H2 with weights 1-gamma, gamma/2, gamma/2
 * (distribute_forces, heatcurrent.f90:541-555); columns of M sites are not summed by
 * compute_heat_flux2 (j <= 3*mxatms/4, heatcurrent.f90:711-715). */
static void fm_entry(const ctx_t *C, int a, int b, const double *F)
{
    ora_elec_args *A = C->A;
    int rows[3], nr;
    double wt[3];
    if (a < C->nreal) {
        rows[0] = a;
        wt[0] = 1.0;
        nr = 1;
    } else {
        const int w = a - C->nreal;
        rows[0] = 3 * w;
        rows[1] = 3 * w + 1;
        rows[2] = 3 * w + 2;
        wt[0] = 1.0 - A->gamma;
        wt[1] = wt[2] = 0.5 * A->gamma;
        nr = 3;
    }
    for (int q = 0; q < nr; ++q) {
        const int r = rows[q];
        if (A->force_matrix) {
            const long n = C->n;
            for (int k = 0; k < 3; ++k) A->force_matrix[(long)k * n * n + (long)b * n + r] += wt[q] * F[k];
        }
        if (!A->heat_w || b >= C->nreal || r == b) continue;
        const double d[3] = {C->x[r] - C->x[b], C->y[r] - C->y[b], C->z[r] - C->z[b]};
        for (int p = 0; p < 3; ++p)
            for (int s = 0; s < 3; ++s) A->heat_w[9 * r + 3 * p + s] += wt[q] * F[p] * d[s];
    }
}

/* Thole athole selection of excluded pairs: athole12 when either site index is
 * 1 mod 3 in the reference's 1-based numbering over the water block, i.e. an oxygen, and
 * also every third M site (rfield.f:439-445, ewald3p.f:196-201; no non-water atoms here). */
static double athole_intra(const ora_elec_args *A, int i, int j)
{
    return (i % 3 == 0 || j % 3 == 0) ? A->athole12 : A->athole13;
}

/* ------------------------------------------------------------------ rfield ---- */

static void rfield(ctx_t *C, int iloop)
{
    ora_elec_args *A = C->A;
    const double *chge = C->chge, *polr = C->polr, *polr2 = C->polr2, *dip = C->dip;
    double *efk = C->efk, *emu = C->emu;
    for (long p = 0; p < C->npair; ++p) {
        const int i = C->pi[p], j = C->pj[p];
        double d[3] = {C->x[i] - C->x[j], C->y[i] - C->y[j], C->z[i] - C->z[j]};
        img(C, d);
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double rr = sqrt(r2);
        if (!(A->rcut > rr)) continue;
        double bbb[4];
        table_interp(C->tab, C->drewd, rr, bbb);
        const double rsqi = 1.0 / r2;
        const double r3i = 1.0 / (rr * rr * rr);
        const double r5i = r3i * rsqi;
        const double bb = pow(polr[i] * polr[j], 1.0 / 6.0);
        if (iloop == 1) {
            double eki = R4PIE0 * chge[j] * (bbb[1] + r3i);
            double ekj = R4PIE0 * chge[i] * (bbb[1] + r3i);
            const double u = rr / bb;
            const double arob4 = A->ascd * (u * u) * (u * u);
            const double dum1 = exp(-arob4);
            eki = eki - R4PIE0 * chge[j] * dum1 * r3i;
            ekj = ekj - R4PIE0 * chge[i] * dum1 * r3i;
            for (int q = 0; q < 3; ++q) {
                efk[3 * i + q] += eki * d[q];
                efk[3 * j + q] -= ekj * d[q];
            }
        }
        if (polr2[i] > 1.e-6 && polr2[j] > 1.e-6) {
            const double rij[6] = {d[0] * d[0], d[0] * d[1], d[0] * d[2], d[1] * d[1], d[1] * d[2], d[2] * d[2]};
            static const int diag[6] = {1, 0, 0, 1, 0, 1};
            const double u = rr / bb;
            const double arob4 = A->athole * (u * u) * (u * u);
            const double dum1 = exp(-arob4);
            const double dum2 = (1.0 + (4.0 / 3.0) * arob4) * dum1;
            double dt[6];
            for (int q = 0; q < 6; ++q) {
                dt[q] = R4PIE0 * ((3.0 * rij[q] * r5i - (diag[q] ? r3i : 0.0)) + rij[q] * bbb[2] -
                                  (diag[q] ? bbb[1] : 0.0));
                dt[q] = dt[q] - (3.0 * rij[q] * r5i * dum2 - (diag[q] ? dum1 * r3i : 0.0)) * R4PIE0;
            }
            const double *mi = dip + 3 * i, *mj = dip + 3 * j;
            emu[3 * i] += dt[0] * mj[0] + dt[1] * mj[1] + dt[2] * mj[2];
            emu[3 * i + 1] += dt[1] * mj[0] + dt[3] * mj[1] + dt[4] * mj[2];
            emu[3 * i + 2] += dt[2] * mj[0] + dt[4] * mj[1] + dt[5] * mj[2];
            emu[3 * j] += dt[0] * mi[0] + dt[1] * mi[1] + dt[2] * mi[2];
            emu[3 * j + 1] += dt[1] * mi[0] + dt[3] * mi[1] + dt[4] * mi[2];
            emu[3 * j + 2] += dt[2] * mi[0] + dt[4] * mi[1] + dt[5] * mi[2];
        }
    }
    /* excluded (intramolecular) pairs: the 6 pairs among O, H1, H2, M of each water */
    for (int w = 0; w < C->nw; ++w) {
        const int s[4] = {3 * w, 3 * w + 1, 3 * w + 2, C->nreal + w};
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) {
                const int i = s[a], j = s[b];
                double d[3] = {C->x[i] - C->x[j], C->y[i] - C->y[j], C->z[i] - C->z[j]};
                img(C, d);
                const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const double rr = sqrt(r2);
                const double rsqi = 1.0 / r2;
                const double r3i = 1.0 / (rr * r2);
                const double r5i = r3i * rsqi;
                double bbb[4];
                bfun_direct(A->alpha, rr, rsqi, bbb);
                if (iloop == 1) {
                    const double eki = R4PIE0 * chge[j] * bbb[1];
                    const double ekj = R4PIE0 * chge[i] * bbb[1];
                    for (int q = 0; q < 3; ++q) {
                        efk[3 * i + q] += eki * d[q];
                        efk[3 * j + q] -= ekj * d[q];
                    }
                }
                if (polr2[i] > 1.e-6 && polr2[j] > 1.e-6) {
                    const double rij[6] = {d[0] * d[0], d[0] * d[1], d[0] * d[2],
                                           d[1] * d[1], d[1] * d[2], d[2] * d[2]};
                    static const int diag[6] = {1, 0, 0, 1, 0, 1};
                    const double bb = pow(polr[i] * polr[j], 1.0 / 6.0);
                    const double u = rr / bb;
                    const double arob4 = athole_intra(A, i, j) * (u * u) * (u * u);
                    const double dum1 = exp(-arob4);
                    const double dum2 = (1.0 + (4.0 / 3.0) * arob4) * dum1;
                    double dt[6];
                    for (int q = 0; q < 6; ++q) {
                        dt[q] = rij[q] * bbb[2] - (diag[q] ? bbb[1] : 0.0);
                        dt[q] = dt[q] + 3.0 * rij[q] * r5i - (diag[q] ? r3i : 0.0);
                        dt[q] = dt[q] - 3.0 * rij[q] * r5i * dum2 + (diag[q] ? dum1 * r3i : 0.0);
                    }
                    double mi[3], mj[3];
                    for (int q = 0; q < 3; ++q) {
                        mi[q] = dip[3 * i + q] * R4PIE0;
                        mj[q] = dip[3 * j + q] * R4PIE0;
                    }
                    emu[3 * i] += dt[0] * mj[0] + dt[1] * mj[1] + dt[2] * mj[2];
                    emu[3 * i + 1] += dt[1] * mj[0] + dt[3] * mj[1] + dt[4] * mj[2];
                    emu[3 * i + 2] += dt[2] * mj[0] + dt[4] * mj[1] + dt[5] * mj[2];
                    emu[3 * j] += dt[0] * mi[0] + dt[1] * mi[1] + dt[2] * mi[2];
                    emu[3 * j + 1] += dt[1] * mi[0] + dt[3] * mi[1] + dt[4] * mi[2];
                    emu[3 * j + 2] += dt[2] * mi[0] + dt[4] * mi[1] + dt[5] * mi[2];
                }
            }
    }
}

/* ----------------------------------------------------------------- ewald1p ---- */

static void recip_setup(ctx_t *C)
{
    ora_elec_args *A = C->A;
    const double *a = A->cell;
    double *b = C->rc;
    b[0] = a[4] * a[8] - a[5] * a[7];
    b[1] = a[2] * a[7] - a[1] * a[8];
    b[2] = a[1] * a[5] - a[2] * a[4];
    b[3] = a[5] * a[6] - a[3] * a[8];
    b[4] = a[0] * a[8] - a[2] * a[6];
    b[5] = a[2] * a[3] - a[0] * a[5];
    b[6] = a[3] * a[7] - a[4] * a[6];
    b[7] = a[1] * a[6] - a[0] * a[7];
    b[8] = a[0] * a[4] - a[1] * a[3];
    const double det = a[0] * b[0] + a[3] * b[1] + a[6] * b[2];
    for (int i = 0; i < 9; ++i) b[i] /= det;
    C->volm = fabs(det);
    /* dcell(rcell): perpendicular widths of the reciprocal cell (buffer(7..9)) */
    const double *r = C->rc;
    const double ax[3] = {r[0], r[1], r[2]}, bx[3] = {r[3], r[4], r[5]}, cx[3] = {r[6], r[7], r[8]};
    const double bc[3] = {bx[1] * cx[2] - bx[2] * cx[1], bx[2] * cx[0] - bx[0] * cx[2], bx[0] * cx[1] - bx[1] * cx[0]};
    const double ca[3] = {cx[1] * ax[2] - cx[2] * ax[1], cx[2] * ax[0] - cx[0] * ax[2], cx[0] * ax[1] - cx[1] * ax[0]};
    const double ab[3] = {ax[1] * bx[2] - ax[2] * bx[1], ax[2] * bx[0] - ax[0] * bx[2], ax[0] * bx[1] - ax[1] * bx[0]};
    const double vol = fabs(ax[0] * bc[0] + ax[1] * bc[1] + ax[2] * bc[2]);
    const double w0 = vol / sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
    const double w1 = vol / sqrt(ca[0] * ca[0] + ca[1] * ca[1] + ca[2] * ca[2]);
    const double w2 = vol / sqrt(ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2]);
    const double twopi = 2.0 * PI_;
    double rcpcut = A->kmax[0] * w0;
    if (A->kmax[1] * w1 < rcpcut) rcpcut = A->kmax[1] * w1;
    if (A->kmax[2] * w2 < rcpcut) rcpcut = A->kmax[2] * w2;
    rcpcut = rcpcut * 1.05 * twopi;
    C->rcpct2 = rcpcut * rcpcut;
    C->rvolm = twopi / C->volm;
    C->ralph = -0.25 / (A->alpha * A->alpha);

    const int n = C->n, km2 = A->kmax[1], km3 = A->kmax[2];
    for (int j = 0; j < n; ++j) {
        const double ssx = r[0] * C->x[j] + r[3] * C->y[j] + r[6] * C->z[j];
        const double ssy = r[1] * C->x[j] + r[4] * C->y[j] + r[7] * C->z[j];
        const double ssz = r[2] * C->x[j] + r[5] * C->y[j] + r[8] * C->z[j];
        C->elc1[j] = cos(twopi * ssx);
        C->els1[j] = sin(twopi * ssx);
        C->emc[j] = 1.0;
        C->ems[j] = 0.0;
        C->enc[j] = 1.0;
        C->ens[j] = 0.0;
        C->emc[n + j] = cos(twopi * ssy);
        C->ems[n + j] = sin(twopi * ssy);
        C->enc[n + j] = cos(twopi * ssz);
        C->ens[n + j] = sin(twopi * ssz);
    }
    for (int l = 2; l <= km2; ++l)
        for (int j = 0; j < n; ++j) {
            C->emc[l * n + j] = C->emc[(l - 1) * n + j] * C->emc[n + j] - C->ems[(l - 1) * n + j] * C->ems[n + j];
            C->ems[l * n + j] = C->ems[(l - 1) * n + j] * C->emc[n + j] + C->emc[(l - 1) * n + j] * C->ems[n + j];
        }
    for (int l = 2; l <= km3; ++l)
        for (int j = 0; j < n; ++j) {
            C->enc[l * n + j] = C->enc[(l - 1) * n + j] * C->enc[n + j] - C->ens[(l - 1) * n + j] * C->ens[n + j];
            C->ens[l * n + j] = C->ens[(l - 1) * n + j] * C->enc[n + j] + C->enc[(l - 1) * n + j] * C->ens[n + j];
        }
}

/* per-atom reciprocal-space stress and energy of one k vector (ewald1p.f:550-693) */
static void ewald1p_atomic(ctx_t *C, const double *k, double akk, double akv2, double Qc, double Qs,
                           const double *Mc, const double *Ms)
{
    ora_elec_args *A = C->A;
    const double scalefactor = 2.0 * C->rvolm * R4PIE0;
    double B[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) B[3 * a + b] = (a == b ? 1.0 : 0.0) - akv2 * k[a] * k[b];
    const double dpkMc = k[0] * Mc[0] + k[1] * Mc[1] + k[2] * Mc[2];
    const double dpkMs = k[0] * Ms[0] + k[1] * Ms[1] + k[2] * Ms[2];
    for (int j = 0; j < C->n; ++j) {
        const double *mu = C->dip + 3 * j;
        const double Cqi = C->chge[j] * C->ckr[j], Sqi = C->chge[j] * C->skr[j];
        const double CMi[3] = {mu[0] * C->ckr[j], mu[1] * C->ckr[j], mu[2] * C->ckr[j]};
        const double SMi[3] = {mu[0] * C->skr[j], mu[1] * C->skr[j], mu[2] * C->skr[j]};
        const double dpkCMi = k[0] * CMi[0] + k[1] * CMi[1] + k[2] * CMi[2];
        const double dpkSMi = k[0] * SMi[0] + k[1] * SMi[1] + k[2] * SMi[2];
        const double cc = Cqi * Qc + Sqi * Qs;
        const double cd = Sqi * dpkMc - Cqi * dpkMs - dpkSMi * Qc + dpkCMi * Qs;
        const double dd = dpkCMi * dpkMc + dpkSMi * dpkMs;
        if (A->s_ew1p)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    const double sCC = B[3 * a + b] * cc;
                    const double sCD = B[3 * a + b] * cd + Sqi * k[a] * Mc[b] - Cqi * k[a] * Ms[b] -
                                       k[a] * SMi[b] * Qc + k[a] * CMi[b] * Qs;
                    const double sDD = B[3 * a + b] * dd + k[a] * CMi[b] * dpkMc + k[a] * SMi[b] * dpkMs +
                                       dpkCMi * k[a] * Mc[b] + dpkSMi * k[a] * Ms[b];
                    A->s_ew1p[9 * j + 3 * a + b] -=
                        sCC * scalefactor * akk + sCD * scalefactor * akk + sDD * scalefactor * akk;
                }
        if (A->e_ew1p) A->e_ew1p[j] += scalefactor * akk * (Cqi * Qc + Sqi * Qs);
    }
}

static void ewald1p(ctx_t *C, int iloop, int iflag)
{
    ora_elec_args *A = C->A;
    const int n = C->n, km1 = A->kmax[0], km2 = A->kmax[1], km3 = A->kmax[2];
    const double twopi = 2.0 * PI_, *rc = C->rc, *chge = C->chge, *polr2 = C->polr2, *dip = C->dip;
    double engc = 0.0, engdum = 0.0, omg[9] = {0};
    if (iflag == 1) {
        C->engsic = 0.0;
        for (int i = 0; i < n; ++i) C->engsic = C->engsic + chge[i] * chge[i];
        C->engsic = -R4PIE0 * A->alpha * C->engsic / SQRPI;
        if (A->e_ew1p)
            for (int i = 0; i < n; ++i) A->e_ew1p[i] += -R4PIE0 * A->alpha * chge[i] * chge[i] / SQRPI;
        memset(C->fk, 0, sizeof(double) * 3 * n);
    }
    for (int j = 0; j < n; ++j) {
        C->elc0[j] = 1.0;
        C->els0[j] = 0.0;
    }
    int mmin = 0, nmin = 1;
    for (int ll = 0; ll <= km1; ++ll) {
        const double tl = twopi * (double)ll;
        const double rkx1 = tl * rc[0], rky1 = tl * rc[3], rkz1 = tl * rc[6];
        if (ll == 1) {
            for (int j = 0; j < n; ++j) {
                C->elc0[j] = C->elc1[j];
                C->els0[j] = C->els1[j];
            }
        } else if (ll > 1) {
            for (int j = 0; j < n; ++j) {
                const double cs = C->elc0[j];
                C->elc0[j] = cs * C->elc1[j] - C->els0[j] * C->els1[j];
                C->els0[j] = C->els0[j] * C->elc1[j] + cs * C->els1[j];
            }
        }
        for (int mm = mmin; mm <= km2; ++mm) {
            const int m = abs(mm);
            const double tm = twopi * (double)mm;
            const double rkx2 = rkx1 + tm * rc[1], rky2 = rky1 + tm * rc[4], rkz2 = rkz1 + tm * rc[7];
            for (int nn = nmin; nn <= km3; ++nn) {
                const int nq = abs(nn);
                const double tn = twopi * (double)nn;
                const double k[3] = {rkx2 + tn * rc[2], rky2 + tn * rc[5], rkz2 + tn * rc[8]};
                const double rksq = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                if (!(rksq <= C->rcpct2)) continue;
                double Mc[3] = {0, 0, 0}, Ms[3] = {0, 0, 0};
                for (int j = 0; j < n; ++j) {
                    const double e0 = C->elc0[j], s0 = C->els0[j];
                    const double em = C->emc[m * n + j], sm = C->ems[m * n + j];
                    double clm, slm;
                    if (mm >= 0) {
                        clm = e0 * em - s0 * sm;
                        slm = s0 * em + sm * e0;
                    } else {
                        clm = e0 * em + s0 * sm;
                        slm = s0 * em - sm * e0;
                    }
                    const double tenc = C->enc[nq * n + j], tens = C->ens[nq * n + j];
                    if (nn >= 0) {
                        C->ckr[j] = clm * tenc - slm * tens;
                        C->skr[j] = slm * tenc + clm * tens;
                    } else {
                        C->ckr[j] = clm * tenc + slm * tens;
                        C->skr[j] = slm * tenc - clm * tens;
                    }
                    if (polr2[j] > 1.e-6)
                        for (int q = 0; q < 3; ++q) {
                            Mc[q] += C->ckr[j] * dip[3 * j + q];
                            Ms[q] += C->skr[j] * dip[3 * j + q];
                        }
                }
                double ckcs = 0.0, ckss = 0.0;
                for (int j = 0; j < n; ++j) {
                    ckcs += chge[j] * C->ckr[j];
                    ckss += chge[j] * C->skr[j];
                }
                const double skdmuckrs = k[0] * Mc[0] + k[1] * Mc[1] + k[2] * Mc[2];
                const double skdmuskrs = k[0] * Ms[0] + k[1] * Ms[1] + k[2] * Ms[2];
                const double summu1 = ckcs - skdmuskrs, summu2 = ckss + skdmuckrs;
                const double rrksq = 1.0 / rksq;
                const double akk = exp(C->ralph * rksq) * rrksq;
                const double akv = 2.0 * akk * (rrksq - C->ralph);
                const double akv2 = 2.0 * (rrksq - C->ralph);
                if (iflag == 1) {
                    ewald1p_atomic(C, k, akk, akv2, ckcs, ckss, Mc, Ms);
                    engc += akk * (ckcs * ckcs + ckss * ckss);
                    engdum += akk * (summu1 * summu1 + summu2 * summu2);
                    const double virprs = akv * (summu1 * summu1 + summu2 * summu2);
                    static const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
                    static const int io[6] = {0, 4, 8, 1, 2, 5};
                    for (int t = 0; t < 6; ++t)
                        omg[io[t]] = omg[io[t]] - virprs * k[ia[t]] * k[ib[t]] -
                                     2.0 * akk * k[ia[t]] * Ms[ib[t]] * summu1 + 2.0 * akk * k[ia[t]] * Mc[ib[t]] * summu2;
                }
                for (int j = 0; j < n; ++j) {
                    double djmuk = 0.0;
                    if (polr2[j] > 1.e-6) djmuk = dip[3 * j] * k[0] + dip[3 * j + 1] * k[1] + dip[3 * j + 2] * k[2];
                    if (iloop == 1) {
                        const double tt = -C->skr[j] * ckcs + C->ckr[j] * ckss;
                        for (int q = 0; q < 3; ++q) C->efdc[3 * j + q] -= k[q] * akk * tt;
                    }
                    if (polr2[j] > 1.e-6 && iflag == 0) {
                        const double vvmu = C->skr[j] * skdmuskrs + C->ckr[j] * skdmuckrs;
                        for (int q = 0; q < 3; ++q) C->efdd[3 * j + q] -= k[q] * akk * vvmu;
                    }
                    if (iflag == 1) {
                        const double ssmu = summu1 * (-C->skr[j] * chge[j] - C->ckr[j] * djmuk) +
                                            summu2 * (-C->skr[j] * djmuk + C->ckr[j] * chge[j]);
                        for (int q = 0; q < 3; ++q) C->fk[3 * j + q] -= k[q] * akk * ssmu;
                        C->vk[j] += akk * (C->ckr[j] * ckcs + C->skr[j] * ckss);
                        C->vdck[j] += akk * (C->skr[j] * skdmuckrs - C->ckr[j] * skdmuskrs);
                    }
                }
            }
            nmin = -km3;
        }
        mmin = -km2;
    }
    const double scale = 4.0 * C->rvolm * R4PIE0;
    for (int j = 0; j < n; ++j) {
        if (iloop == 1)
            for (int q = 0; q < 3; ++q) C->efdc[3 * j + q] = scale * C->efdc[3 * j + q];
        if (iflag == 0) {
            for (int q = 0; q < 3; ++q) C->efdd[3 * j + q] = scale * C->efdd[3 * j + q];
        } else {
            for (int q = 0; q < 3; ++q) C->fk[3 * j + q] = scale * C->fk[3 * j + q];
            C->vk[j] = C->vk[j] * scale;
            C->vdck[j] = C->vdck[j] * scale;
        }
    }
    if (iflag == 0) {
        const double dum = 4.0 * R4PIE0 * A->alpha * A->alpha * A->alpha / SQRPI / 3.0;
        for (int j = 0; j < n; ++j)
            for (int q = 0; q < 3; ++q) C->efdd[3 * j + q] += dum * dip[3 * j + q];
        return;
    }
    for (int j = 0; j < n; ++j) C->vs[j] = -2.0 * R4PIE0 * chge[j] * A->alpha / SQRPI;
    A->eng_ew1p = 2.0 * C->rvolm * R4PIE0 * engc + C->engsic;
    const double scal1 = 2.0 * C->rvolm * R4PIE0;
    double *sk = A->stress_ew1p;
    sk[0] = scal1 * (omg[0] + engdum);
    sk[1] = scal1 * omg[1];
    sk[2] = scal1 * omg[2];
    sk[3] = scal1 * omg[1];
    sk[4] = scal1 * (omg[4] + engdum);
    sk[5] = scal1 * omg[5];
    sk[6] = scal1 * omg[2];
    sk[7] = scal1 * omg[5];
    sk[8] = scal1 * (omg[8] + engdum);
    for (int q = 0; q < 9; ++q) A->stress[q] += sk[q];
    A->vir_ew1p = -scal1 * (omg[0] + omg[4] + omg[8] + 3.0 * engdum);
}

/* ------------------------------------------------------------- pair forces ---- */

static void pair_force_out(ctx_t *C, int i, int j, const double *d, const double *fv, double ffac, double rsq,
                           double virdum)
{
    ora_elec_args *A = C->A;
    A->vir_cpe += -ffac * rsq - virdum;
    for (int q = 0; q < 3; ++q) {
        A->f[3 * i + q] += fv[q];
        A->f[3 * j + q] -= fv[q];
    }
    const double mf[3] = {-fv[0], -fv[1], -fv[2]};
    fm_entry(C, j, i, mf);
    fm_entry(C, i, j, fv);
    A->stress[0] += d[0] * fv[0];
    A->stress[1] += d[0] * fv[1];
    A->stress[2] += d[0] * fv[2];
    A->stress[3] += d[0] * fv[1];
    A->stress[4] += d[1] * fv[1];
    A->stress[5] += d[1] * fv[2];
    A->stress[6] += d[0] * fv[2];
    A->stress[7] += d[1] * fv[2];
    A->stress[8] += d[2] * fv[2];
}

/* ewald2p.f:118-600 for the list pairs (nthole = 4, lacs, lads, no ions) */
static void ewald2p(ctx_t *C)
{
    ora_elec_args *A = C->A;
    const double *chge = C->chge, *polr = C->polr, *polr2 = C->polr2, *dip = C->dip;
    const double g34 = exp(ora_gammln(3.0 / 4.0));
    A->eng_ew2p = 0.0;
    for (long p = 0; p < C->npair; ++p) {
        const int i = C->pi[p], j = C->pj[p];
        double d[3] = {C->x[i] - C->x[j], C->y[i] - C->y[j], C->z[i] - C->z[j]};
        img(C, d);
        const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (!(A->rcut * A->rcut > rsq)) continue;
        const double rrr = sqrt(rsq);
        double bbb[4];
        table_interp(C->tab, C->drewd, rrr, bbb);
        const double rri = 1.0 / rrr, rsqi = 1.0 / rsq;
        const double r3i = 1.0 / (rrr * rrr * rrr);
        const double r5i = r3i * rsqi;
        const double cicj = chge[i] * chge[j];
        const double edum = cicj * (bbb[0] + 1.0 / rrr) * R4PIE0;
        A->eng_ew2p += edum;
        if (A->e_ew2p) {
            A->e_ew2p[i] += 0.5 * edum;
            A->e_ew2p[j] += 0.5 * edum;
        }
        double espdum = (bbb[0] + 1.0 / rrr) * R4PIE0;
        double ffac = cicj * (r3i + bbb[1]) * R4PIE0;
        const double bb = pow(polr[i] * polr[j], 1.0 / 6.0);
        const double u = rrr / bb, u4 = (u * u) * (u * u);
        {   /* smeared charge-charge (ewald2p.f:259-275) */
            const double arob4 = A->ascc * u4;
            const double dum1 = exp(-arob4);
            const double sm = (dum1 - pow(arob4, 0.25) * g34 * ora_gammq(0.75, arob4)) / rrr * R4PIE0;
            A->eng_ew2p -= cicj * sm;
            if (A->e_ew2p) {
                A->e_ew2p[i] += -cicj * 0.5 * sm;
                A->e_ew2p[j] += -cicj * 0.5 * sm;
            }
            espdum -= sm;
            ffac -= cicj * r3i * dum1 * R4PIE0;
        }
        const double *mi = dip + 3 * i, *mj = dip + 3 * j;
        const double rdmui = mi[0] * d[0] + mi[1] * d[1] + mi[2] * d[2];
        const double rdmuj = mj[0] * d[0] + mj[1] * d[1] + mj[2] * d[2];
        const double facmu0 = (chge[j] * rdmui - chge[i] * rdmuj) * (-3.0) * r5i * R4PIE0;
        const double facmui0 = r3i * chge[j] * R4PIE0;
        const double facmuj0 = -r3i * chge[i] * R4PIE0;
        double facmu = facmu0 - (chge[j] * rdmui - chge[i] * rdmuj) * bbb[2] * R4PIE0;
        double dfacmui = facmui0 + bbb[1] * chge[j] * R4PIE0;
        double dfacmuj = facmuj0 - bbb[1] * chge[i] * R4PIE0;
        double espdcdum = (bbb[1] + r3i) * R4PIE0;
        {   /* smeared charge-dipole */
            const double arob4 = A->ascd * u4;
            const double dum1 = exp(-arob4);
            const double dum2 = (1.0 + (4.0 / 3.0) * arob4) * dum1;
            facmu -= facmu0 * dum2;
            dfacmui -= facmui0 * dum1;
            dfacmuj -= facmuj0 * dum1;
            espdcdum -= r3i * dum1 * R4PIE0;
        }
        if (polr2[i] > 1.e-6 && polr2[j] > 1.e-6) {
            const double dpidpj = mi[0] * mj[0] + mi[1] * mj[1] + mi[2] * mj[2];
            facmu += ((dpidpj * 3.0 + rdmui * rdmuj * rsqi * (-15.0)) * r5i + bbb[2] * dpidpj - bbb[3] * rdmui * rdmuj) *
                     R4PIE0;
            dfacmui += (3.0 * r5i + bbb[2]) * rdmuj * R4PIE0;
            dfacmuj += (3.0 * r5i + bbb[2]) * rdmui * R4PIE0;
            const double facmu1 = -15.0 * rdmui * rdmuj * rsqi * r5i * R4PIE0;
            const double facmu2 = 3.0 * dpidpj * r5i * R4PIE0;
            const double dfacmui1 = 3.0 * rdmuj * r5i * R4PIE0;
            const double dfacmuj1 = 3.0 * rdmui * r5i * R4PIE0;
            const double arob4 = A->athole * u4;
            const double dum1 = exp(-arob4);
            const double dum2 = (1.0 + (4.0 / 3.0) * arob4) * dum1;
            const double bb4 = (bb * bb) * (bb * bb);
            const double dum3 = 4.0 * A->athole * (4.0 * arob4 - 1.0) * dum1 * rsqi / bb4;
            const double dum4 = 4.0 * A->athole * dum1 * rrr / bb4;
            facmu = facmu - facmu1 * dum2 - facmu2 * dum1 + rdmui * rdmuj * dum3 * rri * R4PIE0 -
                    dpidpj * rsqi * dum4 * R4PIE0;
            dfacmui -= dfacmui1 * dum2;
            dfacmuj -= dfacmuj1 * dum2;
        }
        const double dfac[3] = {dfacmui * mi[0] + dfacmuj * mj[0], dfacmui * mi[1] + dfacmuj * mj[1],
                                dfacmui * mi[2] + dfacmuj * mj[2]};
        ffac += facmu;
        const double virdum = dfac[0] * d[0] + dfac[1] * d[1] + dfac[2] * d[2];
        const double fv[3] = {ffac * d[0] + dfac[0], ffac * d[1] + dfac[1], ffac * d[2] + dfac[2]};
        pair_force_out(C, i, j, d, fv, ffac, rsq, virdum);
        C->vr[i] += chge[j] * espdum;
        C->vr[j] += chge[i] * espdum;
        C->vdcr[i] += rdmuj * espdcdum;
        C->vdcr[j] -= rdmui * espdcdum;
    }
}

/* ewald3p.f:80-391 over the excluded pairs (nthole = 4) */
static void ewald3p(ctx_t *C)
{
    ora_elec_args *A = C->A;
    const double *chge = C->chge, *polr = C->polr, *polr2 = C->polr2, *dip = C->dip;
    A->eng_ew3p = 0.0;
    for (int w = 0; w < C->nw; ++w) {
        const int s[4] = {3 * w, 3 * w + 1, 3 * w + 2, C->nreal + w};
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) {
                const int i = s[a], j = s[b];
                double d[3] = {C->x[i] - C->x[j], C->y[i] - C->y[j], C->z[i] - C->z[j]};
                img(C, d);
                const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const double rrr = sqrt(rsq);
                const double rri = 1.0 / rrr, rsqi = 1.0 / rsq;
                const double r3i = 1.0 / (rrr * rrr * rrr);
                const double r5i = r3i * rsqi;
                double bbb[4];
                bfun_direct(A->alpha, rrr, rsqi, bbb);
                const double cicj = chge[i] * chge[j];
                A->eng_ew3p += cicj * bbb[0] * R4PIE0;
                if (A->e_ew3p) {
                    A->e_ew3p[i] += 0.5 * cicj * bbb[0] * R4PIE0;
                    A->e_ew3p[j] += 0.5 * cicj * bbb[0] * R4PIE0;
                }
                const double espdum = bbb[0] * R4PIE0;
                const double espdcdum = bbb[1] * R4PIE0;
                double ffac = cicj * bbb[1] * R4PIE0;
                const double *mi = dip + 3 * i, *mj = dip + 3 * j;
                const double rdmui = mi[0] * d[0] + mi[1] * d[1] + mi[2] * d[2];
                const double rdmuj = mj[0] * d[0] + mj[1] * d[1] + mj[2] * d[2];
                double facmu = (chge[j] * rdmui - chge[i] * rdmuj) * (-bbb[2]) * R4PIE0;
                double dfacmui = bbb[1] * chge[j] * R4PIE0;
                double dfacmuj = -bbb[1] * chge[i] * R4PIE0;
                if (polr2[i] > 1.e-6 && polr2[j] > 1.e-6) {
                    const double dpidpj = mi[0] * mj[0] + mi[1] * mj[1] + mi[2] * mj[2];
                    facmu += (bbb[2] * dpidpj - bbb[3] * rdmui * rdmuj) * R4PIE0;
                    dfacmui += bbb[2] * rdmuj * R4PIE0;
                    dfacmuj += bbb[2] * rdmui * R4PIE0;
                    const double facmu1 = -15.0 * rdmui * rdmuj * rsqi * r5i * R4PIE0;
                    const double facmu2 = 3.0 * dpidpj * r5i * R4PIE0;
                    facmu = facmu + facmu1 + facmu2;
                    const double dfacmuidd = 3.0 * rdmuj * r5i * R4PIE0;
                    const double dfacmujdd = 3.0 * rdmui * r5i * R4PIE0;
                    dfacmui += dfacmuidd;
                    dfacmuj += dfacmujdd;
                    const double bb = pow(polr[i] * polr[j], 1.0 / 6.0);
                    const double ath = athole_intra(A, i, j);
                    const double u = rrr / bb;
                    const double arob4 = ath * (u * u) * (u * u);
                    const double dum1 = exp(-arob4);
                    const double dum2 = (1.0 + (4.0 / 3.0) * arob4) * dum1;
                    const double bb4 = (bb * bb) * (bb * bb);
                    const double dum3 = 4.0 * ath * (4.0 * arob4 - 1.0) * dum1 * rsqi / bb4;
                    const double dum4 = 4.0 * ath * dum1 * rrr / bb4;
                    facmu = facmu - facmu1 * dum2 - facmu2 * dum1 + rdmui * rdmuj * dum3 * rri * R4PIE0 -
                            dpidpj * rsqi * dum4 * R4PIE0;
                    dfacmui -= dfacmuidd * dum2;
                    dfacmuj -= dfacmujdd * dum2;
                }
                const double dfac[3] = {dfacmui * mi[0] + dfacmuj * mj[0], dfacmui * mi[1] + dfacmuj * mj[1],
                                        dfacmui * mi[2] + dfacmuj * mj[2]};
                ffac += facmu;
                const double fv[3] = {ffac * d[0] + dfac[0], ffac * d[1] + dfac[1], ffac * d[2] + dfac[2]};
                const double virdum = dfac[0] * d[0] + dfac[1] * d[1] + dfac[2] * d[2];
                pair_force_out(C, i, j, d, fv, ffac, rsq, virdum);
                C->vc[i] += chge[j] * espdum;
                C->vc[j] += chge[i] * espdum;
                C->vc[i] += rdmuj * espdcdum;
                C->vc[j] -= rdmui * espdcdum;
            }
    }
}

/* --------------------------------------------------------------- qdforce ---- */

/* qdforce.f90:70-566: forces (and stress, virial, force-matrix columns) from the
 * position dependence of the DMS charges, f -= sum_site (dq_site/dx) * ESP(site). */
static void qdforce(ctx_t *C)
{
    ora_elec_args *A = C->A;
    const double gamma = A->gamma;
    const double *vesp = A->vesp;
    double strs[6] = {0, 0, 0, 0, 0, 0};
    double qdc[6] = {0, 0, 0, 0, 0, 0};   /* third * vtmp * dqdcos * dcostmp over O, H1, H2 */
    double virqdf = 0.0;
    for (int w = 0; w < C->nw; ++w) {
        const int io = 3 * w, ih1 = io + 1, ih2 = io + 2, im = C->nreal + w;
        const int site[4] = {io, ih1, ih2, im};
        const dms_grad_t *g = &C->dg[w];
        double r1v[3] = {C->x[ih1] - C->x[io], C->y[ih1] - C->y[io], C->z[ih1] - C->z[io]};
        double r2v[3] = {C->x[ih2] - C->x[io], C->y[ih2] - C->y[io], C->z[ih2] - C->z[io]};
        img(C, r1v);
        img(C, r2v);
        const double r1 = sqrt(r1v[0] * r1v[0] + r1v[1] * r1v[1] + r1v[2] * r1v[2]);
        const double r2 = sqrt(r2v[0] * r2v[0] + r2v[1] * r2v[1] + r2v[2] * r2v[2]);
        const double rdot = r1v[0] * r2v[0] + r1v[1] * r2v[1] + r1v[2] * r2v[2];
        const double cost = rdot / (r1 * r2);
        const double r3 = sqrt(r1 * r1 + r2 * r2 - 2.0 * r1 * r2 * cost);
        const double sint = sqrt(1.0 - cost * cost);
        const double cott = cost / sint;
        const double r1sqr = r1 * r1, r2sqr = r2 * r2;
        /* dr[k][site][xyz]: derivative of internal coordinate k (r1, r2, theta) with
         * respect to site (O, H1, H2, M) */
        double dr[3][4][3];
        memset(dr, 0, sizeof dr);
        for (int q = 0; q < 3; ++q) {
            dr[0][1][q] = r1v[q] / r1;
            dr[0][0][q] = -dr[0][1][q];
            dr[1][2][q] = r2v[q] / r2;
            dr[1][0][q] = -dr[1][2][q];
            dr[2][1][q] = cott * (r1v[q] / r1sqr - r2v[q] / rdot);
            dr[2][2][q] = cott * (r2v[q] / r2sqr - r1v[q] / rdot);
            dr[2][0][q] = -dr[2][1][q] - dr[2][2][q];
        }
        double dp1[4][3], dp2[4][3];
        for (int s = 0; s < 4; ++s)
            for (int q = 0; q < 3; ++q) {
                dp1[s][q] = g->dp1dr1 * dr[0][s][q] + g->dp1dr2 * dr[1][s][q] - sint * g->dp1dcos * dr[2][s][q];
                dp2[s][q] = g->dp2dr1 * dr[0][s][q] + g->dp2dr2 * dr[1][s][q] - sint * g->dp2dcos * dr[2][s][q];
            }
        const double tmp = gamma / 2.0 / (1.0 - gamma);
        double dqdr12[4] = {0, 0, 0, 0}, dqdr13[4] = {0, 0, 0, 0}, dqdcos[4] = {0, 0, 0, 0};
        dqdr12[1] = g->dp1dr1 + (g->dp1dr1 + g->dp2dr1) * tmp;
        dqdr13[1] = g->dp1dr2 + (g->dp1dr2 + g->dp2dr2) * tmp;
        dqdcos[1] = g->dp1dcos + (g->dp1dcos + g->dp2dcos) * tmp;
        dqdr12[2] = g->dp2dr1 + (g->dp1dr1 + g->dp2dr1) * tmp;
        dqdr13[2] = g->dp2dr2 + (g->dp1dr2 + g->dp2dr2) * tmp;
        dqdcos[2] = g->dp2dcos + (g->dp1dcos + g->dp2dcos) * tmp;
        dqdr12[3] = -(g->dp1dr1 + g->dp2dr1) / (1.0 - gamma);
        dqdr13[3] = -(g->dp1dr2 + g->dp2dr2) / (1.0 - gamma);
        dqdcos[3] = -(g->dp1dcos + g->dp2dcos) / (1.0 - gamma);
        /* dq[charged site j][moved site s][xyz] */
        double dq[4][4][3];
        memset(dq, 0, sizeof dq);
        for (int s = 0; s < 4; ++s)
            for (int q = 0; q < 3; ++q) {
                dq[1][s][q] = dp1[s][q] * (tmp + 1.0) + dp2[s][q] * tmp;
                dq[2][s][q] = dp2[s][q] * (tmp + 1.0) + dp1[s][q] * tmp;
                dq[3][s][q] = -(dq[1][s][q] + dq[2][s][q]);
            }
        /* forces and force-matrix columns (qdforce.f90:309-348) */
        for (int j = 0; j < 4; ++j) {
            const double vt = vesp[site[j]];
            /* column: the site's own index for O, H1, H2; for M the reference writes
             * column n_water + imol (1-based), i.e. real atom nw + w (0-based). */
            const int col = (j < 3) ? site[j] : (A->compat_qd_column ? C->nw + w : im);
            for (int s = 0; s < 4; ++s) {
                const double F[3] = {-dq[j][s][0] * vt, -dq[j][s][1] * vt, -dq[j][s][2] * vt};
                for (int q = 0; q < 3; ++q) A->f[3 * site[s] + q] += F[q];
                fm_entry(C, site[s], col, F);
            }
        }
        /* stress and virial (qdforce.f90:356-560) */
        double rv[2][3] = {{C->x[io] - C->x[ih1], C->y[io] - C->y[ih1], C->z[io] - C->z[ih1]},
                           {C->x[io] - C->x[ih2], C->y[io] - C->y[ih2], C->z[io] - C->z[ih2]}};
        img(C, rv[0]);
        img(C, rv[1]);
        double r3v[3] = {C->x[ih1] - C->x[ih2], C->y[ih1] - C->y[ih2], C->z[ih1] - C->z[ih2]};
        img(C, r3v);
        const double rr1 = sqrt(rv[0][0] * rv[0][0] + rv[0][1] * rv[0][1] + rv[0][2] * rv[0][2]);
        const double rr2 = sqrt(rv[1][0] * rv[1][0] + rv[1][1] * rv[1][1] + rv[1][2] * rv[1][2]);
        const double rr3 = sqrt(r3v[0] * r3v[0] + r3v[1] * r3v[1] + r3v[2] * r3v[2]);
        const double dcosdr12 = (r1 - r2 * cost) / r1 / r2;
        const double dcosdr13 = (r2 - r1 * cost) / r1 / r2;
        const double dcosdr23 = -r3 / r1 / r2;
        static const int sa[6] = {0, 0, 0, 1, 1, 2}, sb[6] = {0, 1, 2, 1, 2, 2};
        for (int i = 0; i < 4; ++i) {
            const double vt = vesp[site[i]];
            for (int pr = 0; pr < 2; ++pr) {
                const double dqdr = pr == 0 ? dqdr12[i] : dqdr13[i];
                const double *r = rv[pr];
                const double rjk = pr == 0 ? rr1 : rr2;
                for (int t = 0; t < 6; ++t) strs[t] -= vt * dqdr * r[sa[t]] * r[sb[t]] / rjk;
                virqdf += vt * rjk * dqdr;
            }
            double dc[6];
            for (int t = 0; t < 6; ++t)
                dc[t] = dcosdr12 * rv[0][sa[t]] * rv[0][sb[t]] / rr1 + dcosdr13 * rv[1][sa[t]] * rv[1][sb[t]] / rr2 +
                        dcosdr23 * r3v[sa[t]] * r3v[sb[t]] / rr3;
            for (int t = 0; t < 6; ++t) strs[t] -= vt * dqdcos[i] * dc[t];
            for (int t = 0; t < 6; ++t) qdc[t] += vt * dqdcos[i] * dc[t];
            virqdf += vt * (dc[0] + dc[3] + dc[5]) * dqdcos[i];
        }
    }
    {
        const int q9[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
        for (int q = 0; q < 9; ++q) A->stress_qd_atoms[q] = qdc[q9[q]];
    }
    A->stress[0] += strs[0];
    A->stress[1] += strs[1];
    A->stress[2] += strs[2];
    A->stress[3] += strs[1];
    A->stress[4] += strs[3];
    A->stress[5] += strs[4];
    A->stress[6] += strs[2];
    A->stress[7] += strs[4];
    A->stress[8] += strs[5];
    A->vir_qd = virqdf;
}

/* qdforce alone for given site potentials (a test aid, to compare with the reference's own
 * qdforce): A->vesp in, A->f, A->stress, A->force_matrix, A->heat_w accumulated,
 * A->vir_qd and A->stress_qd_atoms set.  Positions as given (the M sites are not rebuilt). */
int ora_qdforce(ora_elec_args *A)
{
    if (A->imcon < 0 || A->imcon > 3) return -1;
    if (!g_dms_loaded) return -3;
    ctx_t C;
    memset(&C, 0, sizeof C);
    C.A = A;
    C.nw = A->nw;
    C.nreal = 3 * A->nw;
    C.n = 4 * A->nw;
    C.x = A->x;
    C.y = A->y;
    C.z = A->z;
    C.dg = calloc(C.nw, sizeof(dms_grad_t));
    if (!C.dg) return -4;
    for (int w = 0; w < C.nw; ++w) {
        const int io = 3 * w;
        const double o[3] = {C.x[io], C.y[io], C.z[io]};
        const double h1[3] = {C.x[io + 1], C.y[io + 1], C.z[io + 1]};
        const double h2[3] = {C.x[io + 2], C.y[io + 2], C.z[io + 2]};
        double qh1, qh2;
        dms_water(A->imcon, A->cell, o, h1, h2, A->gamma, &qh1, &qh2, &C.dg[w], NULL);
    }
    qdforce(&C);
    free(C.dg);
    return 0;
}

/* ------------------------------------------------------------------- driver ---- */

static int elec_forces_run(ora_elec_args *A);

int ora_elec_forces(ora_elec_args *A)
{
    if (A->imcon < 0 || A->imcon > 3) return -1;
    /* imcon 0, an isolated cluster: direct sums over all pairs within rcut, no Ewald split
     * (alpha must be 0) and no reciprocal space */
    if (A->imcon == 0 && A->alpha != 0.0) return -1;
    const int kmax_in[3] = {A->kmax[0], A->kmax[1], A->kmax[2]};
    if (A->imcon == 0) A->kmax[0] = A->kmax[1] = A->kmax[2] = 0;   /* restored on return */
    const int rc_all = elec_forces_run(A);
    for (int q = 0; q < 3; ++q) A->kmax[q] = kmax_in[q];
    return rc_all;
}

static int elec_forces_run(ora_elec_args *A)
{
    if (!g_dms_loaded) return -3;
    ctx_t C;
    memset(&C, 0, sizeof C);
    C.A = A;
    C.nw = A->nw;
    C.nreal = 3 * A->nw;
    C.n = 4 * A->nw;
    const int n = C.n, nreal = C.nreal, nw = C.nw;
    C.x = A->x;
    C.y = A->y;
    C.z = A->z;
    C.chge = A->chge;
    C.dip = A->dip;
    C.polr = calloc(n, sizeof(double));
    C.polr2 = calloc(n, sizeof(double));
    C.dg = calloc(nw, sizeof(dms_grad_t));

    /* M sites (dlpoly_cmd.f:2041-2057) and DMS charges (qdetermine.f90:110-259) */
    for (int w = 0; w < nw; ++w) {
        const int io = 3 * w;
        double d1[3] = {C.x[io + 1] - C.x[io], C.y[io + 1] - C.y[io], C.z[io + 1] - C.z[io]};
        double d2[3] = {C.x[io + 2] - C.x[io], C.y[io + 2] - C.y[io], C.z[io + 2] - C.z[io]};
        img(&C, d1);
        img(&C, d2);
        C.x[nreal + w] = C.x[io] + (d1[0] + d2[0]) * A->gamma / 2.0;
        C.y[nreal + w] = C.y[io] + (d1[1] + d2[1]) * A->gamma / 2.0;
        C.z[nreal + w] = C.z[io] + (d1[2] + d2[2]) * A->gamma / 2.0;
    }
    for (int w = 0; w < nw; ++w) {
        const int io = 3 * w;
        const double o[3] = {C.x[io], C.y[io], C.z[io]};
        const double h1[3] = {C.x[io + 1], C.y[io + 1], C.z[io + 1]};
        const double h2[3] = {C.x[io + 2], C.y[io + 2], C.z[io + 2]};
        double qh1, qh2;
        dms_water(A->imcon, A->cell, o, h1, h2, A->gamma, &qh1, &qh2, &C.dg[w], NULL);
        C.chge[io] = A->q_o;
        C.chge[io + 1] = qh1;
        C.chge[io + 2] = qh2;
        C.chge[nreal + w] = -(qh1 + qh2);
        C.polr[io] = A->polr_site[0];
        C.polr[io + 1] = C.polr[io + 2] = A->polr_site[1];
        C.polr[nreal + w] = A->polr_site[2];
        C.polr2[io] = A->polr2_site[0];
        C.polr2[io + 1] = C.polr2[io + 2] = A->polr2_site[1];
        C.polr2[nreal + w] = A->polr2_site[2];
    }

    /* Verlet pairs: every site pair of different waters; each routine applies its own
     * r < rcut test (rfield.f:92, ewald2p.f:150) */
    long cap = 1024;
    C.pi = malloc(sizeof(int) * cap);
    C.pj = malloc(sizeof(int) * cap);
    const double rl2 = (A->rcut + 0.5) * (A->rcut + 0.5);
    A->npairs = 0;
    for (int i = 0; i < n; ++i) {
        const int wi = i < nreal ? i / 3 : i - nreal;
        for (int j = i + 1; j < n; ++j) {
            const int wj = j < nreal ? j / 3 : j - nreal;
            if (wi == wj) continue;
            double d[3] = {C.x[i] - C.x[j], C.y[i] - C.y[j], C.z[i] - C.z[j]};
            img(&C, d);
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 >= rl2) continue;
            if (A->rcut * A->rcut > r2) ++A->npairs;
            if (C.npair == cap) {
                cap *= 2;
                C.pi = realloc(C.pi, sizeof(int) * cap);
                C.pj = realloc(C.pj, sizeof(int) * cap);
            }
            C.pi[C.npair] = i;
            C.pj[C.npair] = j;
            ++C.npair;
        }
    }

    /* tables and reciprocal-space setup */
    const int mxegrd = ora_erfc_grid(A->rcut, &C.drewd);
    C.tab = malloc(sizeof(double) * 4 * (mxegrd + 1));
    ora_erfc_table(A->alpha, A->rcut, C.tab);
    const int km2 = A->kmax[1], km3 = A->kmax[2];
    C.elc1 = malloc(sizeof(double) * n);
    C.els1 = malloc(sizeof(double) * n);
    C.elc0 = malloc(sizeof(double) * n);
    C.els0 = malloc(sizeof(double) * n);
    C.ckr = malloc(sizeof(double) * n);
    C.skr = malloc(sizeof(double) * n);
    C.emc = malloc(sizeof(double) * n * (km2 + 1));
    C.ems = malloc(sizeof(double) * n * (km2 + 1));
    C.enc = malloc(sizeof(double) * n * (km3 + 1));
    C.ens = malloc(sizeof(double) * n * (km3 + 1));
    C.fk = calloc(3 * n, sizeof(double));
    if (A->imcon != 0) {
        recip_setup(&C);
    } else {
        C.volm = 0.0;
        C.rvolm = 0.0;
        C.ralph = 0.0;
    }
    A->volm = C.volm;

    C.efk = calloc(3 * n, sizeof(double));
    C.emu = calloc(3 * n, sizeof(double));
    C.efdc = calloc(3 * n, sizeof(double));
    C.efdd = calloc(3 * n, sizeof(double));
    C.vk = calloc(n, sizeof(double));
    C.vr = calloc(n, sizeof(double));
    C.vs = calloc(n, sizeof(double));
    C.vc = calloc(n, sizeof(double));
    C.vdck = calloc(n, sizeof(double));
    C.vdcr = calloc(n, sizeof(double));

    /* induced dipoles (forces.f:282-371, iter.f:45-101) */
    int rc = 0;
    int iloop = 1, iflag = 0;
    for (;;) {
        memset(C.efdd, 0, sizeof(double) * 3 * n);
        memset(C.emu, 0, sizeof(double) * 3 * n);
        rfield(&C, iloop);
        ewald1p(&C, iloop, 0);
        const double oldfac = A->omega, fac = 1.0 - oldfac;
        double deltadip = 0.0;
        int ndip = 0;
        for (int i = 0; i < n; ++i) {
            if (!(C.polr2[i] > A->toler)) continue;
            ++ndip;
            double old[3];
            for (int q = 0; q < 3; ++q) {
                old[q] = C.dip[3 * i + q];
                C.dip[3 * i + q] = oldfac * old[q] + fac * C.polr2[i] *
                                                         (C.efk[3 * i + q] + C.emu[3 * i + q] + C.efdd[3 * i + q] +
                                                          C.efdc[3 * i + q] + 0.0) /
                                                         R4PIE0;
            }
            deltadip = deltadip + ((C.dip[3 * i] - old[0]) * (C.dip[3 * i] - old[0]) +
                                   (C.dip[3 * i + 1] - old[1]) * (C.dip[3 * i + 1] - old[1]) +
                                   (C.dip[3 * i + 2] - old[2]) * (C.dip[3 * i + 2] - old[2]));
        }
        if (ndip == 0) ndip = 1;
        deltadip = sqrt(deltadip / (double)ndip);
        A->last_delta = deltadip;
        if (deltadip <= A->toler) iflag = 1;
        ++iloop;
        if (iloop > A->max_iter) {
            rc = -2; /* the reference stops with "dipole diverged!" (forces.f:351-359) */
            break;
        }
        if (iflag) break;
    }
    A->niter = iloop - 1;

    if (rc == 0) {
        /* coulomb_polar.f:58-103 */
        A->eng_polar = 0.0;
        for (int i = 0; i < n; ++i) {
            if (!(C.polr2[i] > 1.e-6)) continue;
            double e = 0.0;
            for (int q = 0; q < 3; ++q)
                e += C.dip[3 * i + q] *
                     (C.efk[3 * i + q] + C.efdc[3 * i + q] + 0.5 * (C.emu[3 * i + q] + C.efdd[3 * i + q]) + 0.0);
            const double edum = C.dip[3 * i] * C.dip[3 * i] + C.dip[3 * i + 1] * C.dip[3 * i + 1] +
                                C.dip[3 * i + 2] * C.dip[3 * i + 2];
            A->eng_polar = A->eng_polar - e;
            A->eng_polar = A->eng_polar + R4PIE0 * edum / C.polr2[i] / 2.0;
            if (A->e_polar) {
                A->e_polar[i] += -e;
                A->e_polar[i] += R4PIE0 * edum / C.polr2[i] / 2.0;
            }
        }
        if (A->efield)
            for (int i = 0; i < 3 * n; ++i) A->efield[i] = C.efk[i] + C.emu[i] + C.efdd[i] + C.efdc[i];
        /* final reciprocal-space pass (forces.f:388-398) */
        A->vir_cpe = 0.0;
        ewald1p(&C, iloop, 1);
        for (int i = 0; i < 3 * n; ++i) A->f[i] += C.fk[i];
        if (A->f_ew1p)
            for (int i = 0; i < 3 * n; ++i) A->f_ew1p[i] += C.fk[i];
        A->vir_cpe += A->vir_ew1p;
        /* real space and excluded pairs (forces.f:585-610, 700-712) */
        ewald2p(&C);
        ewald3p(&C);
        /* ESP and charge-derivative forces (forces.f:874-892) */
        for (int i = 0; i < n; ++i) A->vesp[i] = C.vk[i] + C.vr[i] + C.vs[i] + C.vc[i] + C.vdck[i] + C.vdcr[i] + 0.0;
        qdforce(&C);
        A->vir_cpe += A->vir_qd;
        A->eng_cpe = A->eng_polar + A->eng_ew1p + A->eng_ew2p + A->eng_ew3p;
    }

    free(C.polr); free(C.polr2); free(C.dg); free(C.pi); free(C.pj); free(C.tab);
    free(C.elc1); free(C.els1); free(C.elc0); free(C.els0); free(C.ckr); free(C.skr);
    free(C.emc); free(C.ems); free(C.enc); free(C.ens); free(C.fk);
    free(C.efk); free(C.emu); free(C.efdc); free(C.efdd);
    free(C.vk); free(C.vr); free(C.vs); free(C.vc); free(C.vdck); free(C.vdcr);
    return rc;
}

/* M-site redistribution of site-level arrays onto O, H1, H2 with weights 1-gamma,
 * gamma/2, gamma/2 (forces: dlpoly_cmd.f:2583-2597; per-atom energy and stress:
 * distribute_energy / distribute_stress, heatcurrent.f90:302-459); the M entries are
 * zeroed as the reference does for the energies and stresses.  Any pointer may be NULL. */
void ora_msite_distribute(int nw, double gamma, double *f, double *e, double *s9)
{
    const int nreal = 3 * nw;
    for (int w = 0; w < nw; ++w) {
        const int m = nreal + w, io = 3 * w;
        const double wt[3] = {1.0 - gamma, gamma / 2.0, gamma / 2.0};
        for (int a = 0; a < 3; ++a) {
            if (f)
                for (int q = 0; q < 3; ++q) f[3 * (io + a) + q] += wt[a] * f[3 * m + q];
            if (e) e[io + a] += wt[a] * e[m];
            if (s9)
                for (int q = 0; q < 9; ++q) s9[9 * (io + a) + q] += wt[a] * s9[9 * m + q];
        }
        if (e) e[m] = 0.0;
        if (s9)
            for (int q = 0; q < 9; ++q) s9[9 * m + q] = 0.0;
    }
}
