/*
 * pip_oracle.c -- CPU restatement of the MB-pol 2-body and 3-body polynomial leaves.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The polynomial itself is evaluated as a
 * plain sum over the monomial table (every basis polynomial is a sum of distinct
 * unit-coefficient monomials), which is deliberately a different arithmetic order from
 * both the reference's generated Horner code and the GPU kernels.
 *
 * Restated:
 *   x2b_v6x_p::operator()   mbpol-2b-poly.cpp:358-682 (variables :26-85, extra points
 *                           :118-177, switch :340-354, constants :1846-1866)
 *   x3b_v2x_p::operator()   mbpol-3b-poly.cpp:148-312 (variables/gradients :12-82,
 *                           switch :130-144, constants :1486-1506)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    int nvars, nbasis, nmono;
    int *basis;   /* nmono */
    int *deg;     /* nmono */
    int *vars;    /* nmono*4 */
    double *a;    /* nbasis: fitted coefficients */
} table_t;

static table_t T2, T3;

static int load_table(table_t *t, int nvars, int nbasis, const char *mono, const char *coef)
{
    FILE *f = fopen(mono, "r");
    if (!f) return -1;
    int cap = 65536;
    t->basis = malloc(sizeof(int) * cap);
    t->deg = malloc(sizeof(int) * cap);
    t->vars = malloc(sizeof(int) * 4 * cap);
    t->nmono = 0;
    char line[256];
    while (fgets(line, sizeof line, f)) {
        int v[5], n = sscanf(line, "%d %d %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4]);
        if (n < 2) continue;
        int m = t->nmono++;
        if (m >= cap) { fclose(f); return -2; }
        t->basis[m] = v[0];
        t->deg[m] = n - 1;
        for (int q = 0; q < 4; ++q) t->vars[4 * m + q] = q < n - 1 ? v[q + 1] : 0;
    }
    fclose(f);
    f = fopen(coef, "r");
    if (!f) return -3;
    t->a = malloc(sizeof(double) * nbasis);
    for (int i = 0; i < nbasis; ++i)
        if (fscanf(f, "%lf", &t->a[i]) != 1) { fclose(f); return -4; }
    fclose(f);
    t->nvars = nvars;
    t->nbasis = nbasis;
    return 0;
}

int ora_pip_load(const char *mono2b, const char *coef2b, const char *mono3b, const char *coef3b)
{
    if (T2.nmono == 0 && load_table(&T2, 31, 1153, mono2b, coef2b)) return -1;
    if (T3.nmono == 0 && load_table(&T3, 36, 1163, mono3b, coef3b)) return -2;
    return 0;
}

/* Plain monomial sum, accumulated in x87 extended precision (64-bit significand): the
 * polynomials cancel heavily (sum |terms| / |E| ~ 1e6 on liquid-water clusters), so the
 * oracle keeps extra bits to stay well inside the rounding error of any FP64 evaluation. */
static double poly_eval(const table_t *t, const double *a, const double *x, double *g)
{
    long double e = 0.0L, gl[36];
    for (int j = 0; j < t->nvars; ++j) gl[j] = 0.0L;
    for (int m = 0; m < t->nmono; ++m) {
        const int *v = &t->vars[4 * m];
        const long double c = a[t->basis[m]];
        switch (t->deg[m]) {
        case 1:
            e += c * x[v[0]];
            gl[v[0]] += c;
            break;
        case 2: {
            const long double x0 = x[v[0]], x1 = x[v[1]];
            e += c * x0 * x1;
            gl[v[0]] += c * x1;
            gl[v[1]] += c * x0;
            break;
        }
        case 3: {
            const long double x0 = x[v[0]], x1 = x[v[1]], x2 = x[v[2]];
            const long double c01 = c * x0 * x1;
            e += c01 * x2;
            gl[v[0]] += c * x1 * x2;
            gl[v[1]] += c * x0 * x2;
            gl[v[2]] += c01;
            break;
        }
        default: {
            const long double x0 = x[v[0]], x1 = x[v[1]], x2 = x[v[2]], x3 = x[v[3]];
            const long double c01 = c * x0 * x1, x23 = x2 * x3;
            e += c01 * x23;
            gl[v[0]] += c * x1 * x23;
            gl[v[1]] += c * x0 * x23;
            gl[v[2]] += c01 * x3;
            gl[v[3]] += c01 * x2;
            break;
        }
        }
    }
    for (int j = 0; j < t->nvars; ++j) g[j] = (double)gl[j];
    return (double)e;
}

/* sum over monomials of |a_b * monomial(x)|: the scale of the cancellation in E(x) */
double ora_poly_scale(int which, const double *x)
{
    const table_t *t = which == 2 ? &T2 : &T3;
    long double s = 0.0L;
    for (int m = 0; m < t->nmono; ++m) {
        long double p = t->a[t->basis[m]];
        for (int q = 0; q < t->deg[m]; ++q) p *= x[t->vars[4 * m + q]];
        s += p < 0 ? -p : p;
    }
    return (double)s;
}

double ora_poly_eval(int which, const double *a, const double *x, double *g)
{
    const table_t *t = which == 2 ? &T2 : &T3;
    return poly_eval(t, a ? a : t->a, x, g);
}

/* ---------------------------------------------------------------- 2-body ---- */

/* model constants, mbpol-2b-poly.cpp:1848-1863 */
static const double k_HH_intra = -6.480884773303821e-01;
static const double k_OH_intra = 1.674518993682975e+00;
static const double k_HH_coul = 1.148231864355956e+00;
static const double k_OH_coul = 1.205989761123099e+00;
static const double k_OO_coul = 1.395357065790959e+00;
static const double k_XH_main = 7.347036852042255e-01;
static const double k_XO_main = 7.998249864422826e-01;
static const double k_XX_main = 7.960663960630585e-01;
static const double in_plane_gamma = -9.721486914088159e-02;
static const double out_of_plane_gamma = 9.859272078406150e-02;
static const double r2i = 4.5, r2f = 6.5;

/* site offsets in the 30-coordinate cluster array (mbpol-2b-poly.cpp:376-389) */
enum { Oa = 0, Ha1 = 3, Ha2 = 6, Ob = 9, Hb1 = 12, Hb2 = 15, Xa1 = 18, Xa2 = 21, Xb1 = 24, Xb2 = 27 };

/* one variable: kind 0 = exp(k(d0-r)), kind 1 = exp(k(d0-r))/r  (:26-70) */
typedef struct { int kind; double d0; const double *k; int o1, o2; } var2b_t;

static const var2b_t V2[31] = {
    {0, 1.0, &k_HH_intra, Ha1, Ha2}, {0, 1.0, &k_HH_intra, Hb1, Hb2},
    {0, 1.0, &k_OH_intra, Oa, Ha1},  {0, 1.0, &k_OH_intra, Oa, Ha2},
    {0, 1.0, &k_OH_intra, Ob, Hb1},  {0, 1.0, &k_OH_intra, Ob, Hb2},
    {1, 4.0, &k_HH_coul, Ha1, Hb1},  {1, 4.0, &k_HH_coul, Ha1, Hb2},
    {1, 4.0, &k_HH_coul, Ha2, Hb1},  {1, 4.0, &k_HH_coul, Ha2, Hb2},
    {1, 4.0, &k_OH_coul, Oa, Hb1},   {1, 4.0, &k_OH_coul, Oa, Hb2},
    {1, 4.0, &k_OH_coul, Ob, Ha1},   {1, 4.0, &k_OH_coul, Ob, Ha2},
    {1, 4.0, &k_OO_coul, Oa, Ob},
    {0, 4.0, &k_XH_main, Xa1, Hb1},  {0, 4.0, &k_XH_main, Xa1, Hb2},
    {0, 4.0, &k_XH_main, Xa2, Hb1},  {0, 4.0, &k_XH_main, Xa2, Hb2},
    {0, 4.0, &k_XH_main, Xb1, Ha1},  {0, 4.0, &k_XH_main, Xb1, Ha2},
    {0, 4.0, &k_XH_main, Xb2, Ha1},  {0, 4.0, &k_XH_main, Xb2, Ha2},
    {0, 4.0, &k_XO_main, Oa, Xb1},   {0, 4.0, &k_XO_main, Oa, Xb2},
    {0, 4.0, &k_XO_main, Ob, Xa1},   {0, 4.0, &k_XO_main, Ob, Xa2},
    {0, 4.0, &k_XX_main, Xa1, Xb1},  {0, 4.0, &k_XX_main, Xa1, Xb2},
    {0, 4.0, &k_XX_main, Xa2, Xb1},  {0, 4.0, &k_XX_main, Xa2, Xb2},
};

/* monomer::setup (mbpol-2b-poly.cpp:118-140): two extra points per water */
static void extra_points(const double *ohh, double *oh1, double *oh2, double *x1, double *x2)
{
    for (int i = 0; i < 3; ++i) {
        oh1[i] = ohh[i + 3] - ohh[i];
        oh2[i] = ohh[i + 6] - ohh[i];
    }
    const double v[3] = {oh1[1] * oh2[2] - oh1[2] * oh2[1],
                         oh1[2] * oh2[0] - oh1[0] * oh2[2],
                         oh1[0] * oh2[1] - oh1[1] * oh2[0]};
    for (int i = 0; i < 3; ++i) {
        const double in_plane = ohh[i] + 0.5 * in_plane_gamma * (oh1[i] + oh2[i]);
        const double out_of_plane = out_of_plane_gamma * v[i];
        x1[i] = in_plane + out_of_plane;
        x2[i] = in_plane - out_of_plane;
    }
}

/* monomer::grads (mbpol-2b-poly.cpp:144-176): extra-point gradients to O,H1,H2 */
static void extra_point_grads(const double *oh1, const double *oh2,
                              const double *g1, const double *g2, double *grd)
{
    const double gm[3] = {g1[0] - g2[0], g1[1] - g2[1], g1[2] - g2[2]};
    const double t1[3] = {oh2[1] * gm[2] - oh2[2] * gm[1],
                          oh2[2] * gm[0] - oh2[0] * gm[2],
                          oh2[0] * gm[1] - oh2[1] * gm[0]};
    const double t2[3] = {oh1[1] * gm[2] - oh1[2] * gm[1],
                          oh1[2] * gm[0] - oh1[0] * gm[2],
                          oh1[0] * gm[1] - oh1[1] * gm[0]};
    for (int i = 0; i < 3; ++i) {
        const double gsum = g1[i] + g2[i];
        const double in_plane = 0.5 * in_plane_gamma * gsum;
        const double gh1 = in_plane + out_of_plane_gamma * t1[i];
        const double gh2 = in_plane - out_of_plane_gamma * t2[i];
        grd[i + 0] += gsum - (gh1 + gh2);
        grd[i + 3] += gh1;
        grd[i + 6] += gh2;
    }
}

static void cluster_2b(const double *w1, const double *w2, double *xcrd,
                       double *oha1, double *oha2, double *ohb1, double *ohb2)
{
    memcpy(xcrd, w1, 9 * sizeof(double));
    memcpy(xcrd + 9, w2, 9 * sizeof(double));
    extra_points(xcrd + Oa, oha1, oha2, xcrd + Xa1, xcrd + Xa2);
    extra_points(xcrd + Ob, ohb1, ohb2, xcrd + Xb1, xcrd + Xb2);
}

static double var2b(const var2b_t *v, const double *xcrd, double *dg)
{
    double d[3];
    for (int i = 0; i < 3; ++i) d[i] = xcrd[v->o1 + i] - xcrd[v->o2 + i];
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double k = *v->k;
    const double e = exp(k * (v->d0 - r));
    double val, gg;
    if (v->kind == 0) {
        val = e;
        gg = -k * e / r;
    } else {
        const double rinv = 1.0 / r;
        val = e * rinv;
        gg = -(k + rinv) * val * rinv;
    }
    for (int i = 0; i < 3; ++i) dg[i] = d[i] * gg;
    return val;
}

void ora_vars_2b(const double *w1, const double *w2, double *v31)
{
    double xcrd[30], a1[3], a2[3], b1[3], b2[3], dg[3];
    cluster_2b(w1, w2, xcrd, a1, a2, b1, b2);
    for (int n = 0; n < 31; ++n) v31[n] = var2b(&V2[n], xcrd, dg);
}

/* f_switch (mbpol-2b-poly.cpp:340-354, mbpol-3b-poly.cpp:130-144) */
static double f_switch(double r, double ri, double rf, double *g)
{
    if (r > rf) {
        *g = 0.0;
        return 0.0;
    } else if (r > ri) {
        const double t1 = M_PI / (rf - ri);
        const double x = (r - ri) * t1;
        *g = -sin(x) * t1 / 2.0;
        return (1.0 + cos(x)) / 2.0;
    }
    *g = 0.0;
    return 1.0;
}

double ora_mbpol_2b(const double *w1, const double *w2, double *g1, double *g2)
{
    const double dOO[3] = {w1[0] - w2[0], w1[1] - w2[1], w1[2] - w2[2]};
    const double rOO = sqrt(dOO[0] * dOO[0] + dOO[1] * dOO[1] + dOO[2] * dOO[2]);
    if (rOO > r2f) return 0.0;

    double xcrd[30], oha1[3], oha2[3], ohb1[3], ohb2[3];
    cluster_2b(w1, w2, xcrd, oha1, oha2, ohb1, ohb2);
    double v[31], dv[31][3];
    for (int n = 0; n < 31; ++n) v[n] = var2b(&V2[n], xcrd, dv[n]);
    double g[31];
    const double E_poly = poly_eval(&T2, T2.a, v, g);

    double xgrd[30];
    memset(xgrd, 0, sizeof xgrd);
    for (int n = 0; n < 31; ++n)
        for (int i = 0; i < 3; ++i) {
            const double d = g[n] * dv[n][i];
            xgrd[V2[n].o1 + i] += d;
            xgrd[V2[n].o2 + i] -= d;
        }
    extra_point_grads(oha1, oha2, xgrd + Xa1, xgrd + Xa2, xgrd + Oa);
    extra_point_grads(ohb1, ohb2, xgrd + Xb1, xgrd + Xb2, xgrd + Ob);

    double gsw;
    const double sw = f_switch(rOO, r2i, r2f, &gsw);
    for (int i = 0; i < 9; ++i) {
        g1[i] = sw * xgrd[i];
        g2[i] = sw * xgrd[i + 9];
    }
    gsw *= E_poly / rOO;
    for (int i = 0; i < 3; ++i) {
        const double d = gsw * dOO[i];
        g1[i] += d;
        g2[i] -= d;
    }
    return sw * E_poly;
}

/* ---------------------------------------------------------------- 3-body ---- */

/* model constants, mbpol-3b-poly.cpp:1488-1504 */
static const double r3i = 0.0, r3f = 4.5;
static const double kHH_intra = -2.254303257872797e+00;
static const double kOH_intra = -2.564771901404151e-01;
static const double kHH = 4.247920544074333e-01;
static const double kOH = 8.128985941165371e-01;
static const double kOO = 3.749457984616480e-02;
static const double dHH_intra = 1.690594510379166e+00;
static const double dOH_intra = -2.999999868517452e+00;
static const double dHH = 3.499031358429095e+00;
static const double dOH = 4.854042963514281e+00;
static const double dOO = 4.816312044947604e-08;

/* atom offsets in the 27-coordinate trimer array */
enum { A_O = 0, A_H1 = 3, A_H2 = 6, B_O = 9, B_H1 = 12, B_H2 = 15, C_O = 18, C_H1 = 21, C_H2 = 24 };

typedef struct { const double *k, *r0; int o1, o2; } var3b_t;

/* the 36 variables (mbpol-3b-poly.cpp:170-206) */
static const var3b_t V3[36] = {
    {&kHH_intra, &dHH_intra, A_H1, A_H2}, {&kHH_intra, &dHH_intra, B_H1, B_H2},
    {&kHH_intra, &dHH_intra, C_H1, C_H2}, {&kOH_intra, &dOH_intra, A_O, A_H1},
    {&kOH_intra, &dOH_intra, A_O, A_H2},  {&kOH_intra, &dOH_intra, B_O, B_H1},
    {&kOH_intra, &dOH_intra, B_O, B_H2},  {&kOH_intra, &dOH_intra, C_O, C_H1},
    {&kOH_intra, &dOH_intra, C_O, C_H2},
    {&kHH, &dHH, A_H1, B_H1}, {&kHH, &dHH, A_H1, B_H2}, {&kHH, &dHH, A_H1, C_H1},
    {&kHH, &dHH, A_H1, C_H2}, {&kHH, &dHH, A_H2, B_H1}, {&kHH, &dHH, A_H2, B_H2},
    {&kHH, &dHH, A_H2, C_H1}, {&kHH, &dHH, A_H2, C_H2}, {&kHH, &dHH, B_H1, C_H1},
    {&kHH, &dHH, B_H1, C_H2}, {&kHH, &dHH, B_H2, C_H1}, {&kHH, &dHH, B_H2, C_H2},
    {&kOH, &dOH, A_O, B_H1},  {&kOH, &dOH, A_O, B_H2},  {&kOH, &dOH, A_O, C_H1},
    {&kOH, &dOH, A_O, C_H2},  {&kOH, &dOH, B_O, A_H1},  {&kOH, &dOH, B_O, A_H2},
    {&kOH, &dOH, B_O, C_H1},  {&kOH, &dOH, B_O, C_H2},  {&kOH, &dOH, C_O, A_H1},
    {&kOH, &dOH, C_O, A_H2},  {&kOH, &dOH, C_O, B_H1},  {&kOH, &dOH, C_O, B_H2},
    {&kOO, &dOO, A_O, B_O},   {&kOO, &dOO, A_O, C_O},   {&kOO, &dOO, B_O, C_O},
};

static double dist3(const double *p, const double *q, double *dx)
{
    for (int n = 0; n < 3; ++n) dx[n] = p[n] - q[n];
    return sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
}

void ora_vars_3b(const double *w1, const double *w2, const double *w3, double *v36)
{
    double xc[27], dx[3];
    memcpy(xc, w1, 9 * sizeof(double));
    memcpy(xc + 9, w2, 9 * sizeof(double));
    memcpy(xc + 18, w3, 9 * sizeof(double));
    for (int n = 0; n < 36; ++n) {
        const double d = dist3(xc + V3[n].o1, xc + V3[n].o2, dx);
        v36[n] = exp(-*V3[n].k * (d - *V3[n].r0));
    }
}

double ora_mbpol_3b(const double *w1, const double *w2, const double *w3,
                    double *g1, double *g2, double *g3)
{
    double xc[27], gc[27], dx[3];
    memcpy(xc, w1, 9 * sizeof(double));
    memcpy(xc + 9, w2, 9 * sizeof(double));
    memcpy(xc + 18, w3, 9 * sizeof(double));
    double x[36], g[36];
    ora_vars_3b(w1, w2, w3, x);
    double retval = poly_eval(&T3, T3.a, x, g);

    double rab[3], rac[3], rbc[3];
    const double drab = dist3(w1, w2, rab);
    const double drac = dist3(w1, w3, rac);
    const double drbc = dist3(w2, w3, rbc);
    double gab, gac, gbc;
    const double sab = f_switch(drab, r3i, r3f, &gab);
    const double sac = f_switch(drac, r3i, r3f, &gac);
    const double sbc = f_switch(drbc, r3i, r3f, &gbc);
    const double s = sab * sac + sab * sbc + sac * sbc;
    for (int n = 0; n < 36; ++n) g[n] *= s;

    memset(gc, 0, sizeof gc);
    for (int n = 0; n < 36; ++n) {
        const double d = dist3(xc + V3[n].o1, xc + V3[n].o2, dx);
        const double k = *V3[n].k;
        const double gg = -k * g[n] * exp(-k * (d - *V3[n].r0)) / d;
        for (int i = 0; i < 3; ++i) {
            gc[V3[n].o1 + i] += gg * dx[i];
            gc[V3[n].o2 + i] -= gg * dx[i];
        }
    }
    gab *= (sac + sbc) * retval / drab;
    gac *= (sab + sbc) * retval / drac;
    gbc *= (sab + sac) * retval / drbc;
    retval *= s;
    for (int n = 0; n < 3; ++n) {
        gc[A_O + n] += gab * rab[n] + gac * rac[n];
        gc[B_O + n] += gbc * rbc[n] - gab * rab[n];
        gc[C_O + n] -= gac * rac[n] + gbc * rbc[n];
    }
    for (int i = 0; i < 9; ++i) {
        g1[i] += gc[i];
        g2[i] += gc[9 + i];
        g3[i] += gc[18 + i];
    }
    return retval;
}
