/*
 * water_oracle.c -- CPU restatement of the terms around the MB-pol polynomials and the
 * electrostatics that a force evaluation of MB-pol water in the reference also computes:
 * the Partridge-Schwenke monomer energy, the Tang-Toennies-damped dispersion and its
 * long-range correction.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restated (reference file:line):
 *   monomer bonds (mors, expo,    bndfrc.f:67-108 (vb1/gb1, vb2/gb2, vb7/gb7), pair loop
 *   harm)                         :136-330 (per-atom halves, force matrix)
 *   PS coefficients               init_ps_type_pes (module.f90:513-600)
 *   PS monomer energy / forces    ps_type_couple.f90:15-274 through the keya = 9 branch of
 *                                 angfrc.f:160-545 (per-atom energy pterm/3, force matrix)
 *   TT damping                    tang-toennies.f90:5-74 (factorial as written there)
 *   dispersion table              forgen.f:90-100, 262-269, vv12 / gg12 :338-357,
 *                                 grid parset.f:210-215
 *   dispersion pair forces        srfrce2.f:60-281
 *   long-range correction         lrcorrect.f:100-209, el12 / pl12 :270-290
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define PI_ 3.141592653589793

/* ------------------------------------------------------------------- PS table ---- */

static int g_ps_loaded = 0;
static int g_ps_idx[245][3], g_ps_idxm[9][3];
static double g_c5z[245], g_cmass[9];

int ora_ps_load(const char *table, const char *mass)
{
    double cbasis[245], ccore[245], crest[245], cmass0[9];
    FILE *f = fopen(table, "r");
    if (!f) return -1;
    for (int j = 0; j < 245; ++j)
        if (fscanf(f, "%d %d %d %lf %lf %lf %lf", &g_ps_idx[j][0], &g_ps_idx[j][1], &g_ps_idx[j][2], &g_c5z[j],
                   &cbasis[j], &ccore[j], &crest[j]) != 7) {
            fclose(f);
            return -2;
        }
    fclose(f);
    f = fopen(mass, "r");
    if (!f) return -1;
    for (int j = 0; j < 9; ++j)
        if (fscanf(f, "%d %d %d %lf", &g_ps_idxm[j][0], &g_ps_idxm[j][1], &g_ps_idxm[j][2], &cmass0[j]) != 4) {
            fclose(f);
            return -2;
        }
    fclose(f);
    /* init_ps_type_pes (module.f90:513-600) */
    const double hhmass = 1836.152697, hdmass = 3670.483031;
    const double engcnv = 1.19627, f5z = 0.99967788500000, fbasis = 0.15860145369897;
    const double fcore = -1.6351695982132, frest = 1.0;
    for (int i = 0; i < 9; ++i) {
        g_cmass[i] = cmass0[i] / (1.0 / hdmass - 1.0 / hhmass);
        double fac_corr = g_cmass[i] / hhmass;
        if (g_ps_idxm[i][0] == g_ps_idxm[i][1]) fac_corr = fac_corr * 0.5;
        for (int j = 0; j < 245; ++j)
            if (g_ps_idx[j][0] == g_ps_idxm[i][0] && g_ps_idx[j][1] == g_ps_idxm[i][1] &&
                g_ps_idx[j][2] == g_ps_idxm[i][2]) {
                crest[j] = crest[j] - fac_corr;
                break;
            }
        for (int j = 0; j < 245; ++j)
            if (g_ps_idx[j][1] == g_ps_idxm[i][0] && g_ps_idx[j][0] == g_ps_idxm[i][1] &&
                g_ps_idx[j][2] == g_ps_idxm[i][2]) {
                crest[j] = crest[j] - fac_corr;
                break;
            }
    }
    for (int i = 0; i < 245; ++i) g_c5z[i] = f5z * g_c5z[i] + fbasis * cbasis[i] + fcore * ccore[i] + frest * crest[i];
    for (int i = 0; i < 245; ++i) g_c5z[i] = g_c5z[i] * engcnv;
    for (int i = 0; i < 9; ++i) g_cmass[i] = g_cmass[i] * frest * engcnv;
    g_c5z[0] = g_c5z[0] * 2.0;
    g_ps_loaded = 1;
    return 0;
}

/* processed coefficients (for the GPU library's own table check) */
void ora_ps_coefficients(double *c5z, double *cmass)
{
    memcpy(c5z, g_c5z, sizeof g_c5z);
    memcpy(cmass, g_cmass, sizeof g_cmass);
}

/* ps_type_couple.f90:15-274: (roh1, unit vector 1), (roh2, unit vector 2), angle theta
 * -> pterm (energy), fdum = forces on the two H (fdum[0..2] on atom a, [3..5] on atom c),
 * vterm (virial). */
void ora_ps_couple(double roh1, const double *u1, double roh2, const double *u2, double theta, double *pterm,
                   double *fdum, double *vterm)
{
    const double reoh = 0.958649, thetae = 104.3475 * PI_ / 180.0, b1 = 2.0;
    const double cose = cos(thetae);
    const double xm1 = 1.0 / 1836.152697, xm2 = 1.0 / 1836.152697;
    const double cost = cos(theta);
    const double x1 = (roh1 - reoh) / reoh, x2 = (roh2 - reoh) / reoh, x3 = cost - cose;
    double fm[16][3];
    memset(fm, 0, sizeof fm);
    fm[1][0] = fm[1][1] = fm[1][2] = 1.0;
    for (int j = 2; j <= 15; ++j) {
        fm[j][0] = fm[j - 1][0] * x1;
        fm[j][1] = fm[j - 1][1] * x2;
        fm[j][2] = fm[j - 1][2] * x3;
    }
    double pt = 0.0, vt = 0.0, fa[3] = {0, 0, 0}, fc[3] = {0, 0, 0};
    for (int j = 1; j < 245; ++j) {
        const int i1 = g_ps_idx[j][0], i2 = g_ps_idx[j][1], i3 = g_ps_idx[j][2];
        const double c = g_c5z[j];
        const double dum = fm[i1][0] * fm[i2][1] + fm[i2][0] * fm[i1][1];
        pt = pt + c * dum * fm[i3][2];
        for (int q = 0; q < 3; ++q) {
            const double d1 = (i1 - 1) * fm[i1 - 1][0] * u1[q] / reoh;
            const double d2 = (i2 - 1) * fm[i2 - 1][0] * u1[q] / reoh;
            const double d3 = (i3 - 1) * fm[i3 - 1][2] * (u2[q] - cost * u1[q]) / roh1;
            const double dd = d1 * fm[i2][1] * fm[i3][2] + d2 * fm[i1][1] * fm[i3][2] + dum * d3;
            fa[q] = fa[q] - c * dd;
        }
        for (int q = 0; q < 3; ++q) {
            const double d1 = (i2 - 1) * fm[i2 - 1][1] * u2[q] / reoh;
            const double d2 = (i1 - 1) * fm[i1 - 1][1] * u2[q] / reoh;
            const double d3 = (i3 - 1) * fm[i3 - 1][2] * (u1[q] - cost * u2[q]) / roh2;
            const double dd = fm[i1][0] * d1 * fm[i3][2] + fm[i2][0] * d2 * fm[i3][2] + dum * d3;
            fc[q] = fc[q] - c * dd;
        }
        vt = vt + c * (roh1 * fm[i2][1] * (i1 - 1) * fm[i1 - 1][0] / reoh +
                       roh2 * fm[i1][0] * (i2 - 1) * fm[i2 - 1][1] / reoh +
                       roh1 * fm[i1][1] * (i2 - 1) * fm[i2 - 1][0] / reoh +
                       roh2 * fm[i2][0] * (i1 - 1) * fm[i1 - 1][1] / reoh) *
                      fm[i3][2];
    }
    double v1 = 0.0, v2 = 0.0;
    for (int j = 0; j < 9; ++j) {
        const int i1 = g_ps_idxm[j][0], i2 = g_ps_idxm[j][1], i3 = g_ps_idxm[j][2];
        const double c = g_cmass[j];
        v1 = v1 + c * fm[i1][0] * fm[i2][1] * fm[i3][2];
        v2 = v2 + c * fm[i2][0] * fm[i1][1] * fm[i3][2];
        for (int q = 0; q < 3; ++q) {
            const double d1 = (i1 - 1) * fm[i1 - 1][0] * u1[q] / reoh;
            const double d2 = (i2 - 1) * fm[i2 - 1][0] * u1[q] / reoh;
            const double d3 = (i3 - 1) * fm[i3 - 1][2] * (u2[q] - cost * u1[q]) / roh1;
            const double corr = xm1 * (d1 * fm[i2][1] * fm[i3][2] + fm[i1][0] * fm[i2][1] * d3) +
                                xm2 * (d2 * fm[i1][1] * fm[i3][2] + fm[i2][0] * fm[i1][1] * d3);
            fa[q] = fa[q] - c * corr;
        }
        for (int q = 0; q < 3; ++q) {
            const double d1 = (i2 - 1) * fm[i2 - 1][1] * u2[q] / reoh;
            const double d2 = (i1 - 1) * fm[i1 - 1][1] * u2[q] / reoh;
            const double d3 = (i3 - 1) * fm[i3 - 1][2] * (u1[q] - cost * u2[q]) / roh2;
            const double corr = xm1 * (fm[i1][0] * d1 * fm[i3][2] + fm[i1][0] * fm[i2][1] * d3) +
                                xm2 * (fm[i2][0] * d2 * fm[i3][2] + fm[i2][0] * fm[i1][1] * d3);
            fc[q] = fc[q] - c * corr;
        }
        vt = vt + c * (xm1 * roh1 * fm[i2][1] * (i1 - 1) * fm[i1 - 1][0] / reoh +
                       xm1 * roh2 * fm[i1][0] * (i2 - 1) * fm[i2 - 1][1] / reoh +
                       xm2 * roh1 * fm[i1][1] * (i2 - 1) * fm[i2 - 1][0] / reoh +
                       xm2 * roh2 * fm[i2][0] * (i1 - 1) * fm[i1 - 1][1] / reoh) *
                      fm[i3][2];
    }
    pt = pt + xm1 * v1 + xm2 * v2;
    const double efac = exp(-b1 * ((roh1 - reoh) * (roh1 - reoh) + (roh2 - reoh) * (roh2 - reoh)));
    const double term = pt;
    pt = term * efac + g_c5z[0];
    pt = pt + 2.0384848702093006e-6 * 2625.50 / 100.0;
    for (int q = 0; q < 3; ++q) {
        fa[q] = fa[q] * efac + 2.0 * b1 * (roh1 - reoh) * u1[q] * efac * term;
        fc[q] = fc[q] * efac + 2.0 * b1 * (roh2 - reoh) * u2[q] * efac * term;
    }
    const double dummy1 = -2.0 * b1 * (roh1 - reoh) * roh1 * efac;
    const double dummy2 = -2.0 * b1 * (roh2 - reoh) * roh2 * efac;
    vt = dummy1 * term + dummy2 * term + efac * vt;
    *pterm = pt;
    *vterm = vt;
    for (int q = 0; q < 3; ++q) {
        fdum[q] = fa[q];
        fdum[3 + q] = fc[q];
    }
}

/* ------------------------------------------------------------------ dispersion ---- */

/* tang-toennies.f90:5-31 */
double ora_tang_toennies(int n, double x)
{
    double sum = 1.0;
    for (int nn = n; nn >= 1; --nn) sum = 1.0 + sum * x / nn;
    double tt = 1.0 - sum * exp(-x);
    if (fabs(tt) < 1.0e-8) {
        double term = 1.0;
        for (int nn = n; nn >= 1; --nn) term = term * x / nn;
        sum = 0.0;
        for (int nn = n + 1; nn <= 1000; ++nn) {
            term = term * x / nn;
            sum = sum + term;
            if (fabs(term / sum) < 1.0e-8) break;
        }
        tt = sum * exp(-x);
    }
    return tt;
}

/* tang-toennies.f90:33-74: the "factorial" there multiplies z by itself z times */
static double tt_lr(int n, double x)
{
    double fact = 1.0;
    for (int i = n; i >= 1; --i) fact = fact * n;
    double r = ora_tang_toennies(n, x) / pow(x, n - 3) + exp(-x) * (x * x * (3.0 + x) + 6 * (1.0 + x)) / fact;
    return r / (n - 3);
}

/* forgen.f:338-357 */
static double vv12(double r, double c6, double a6, double c8, double a8)
{
    return -c6 * ora_tang_toennies(6, a6 * r) / pow(r, 6) - c8 * ora_tang_toennies(8, a8 * r) / pow(r, 8);
}

static double gg12(double r, double c6, double a6, double c8, double a8)
{
    return -6 * c6 * ora_tang_toennies(6, a6 * r) / pow(r, 6) - 8 * c8 * ora_tang_toennies(8, a8 * r) / pow(r, 8) +
           r * c6 * pow(a6, 7) * exp(-a6 * r) / 720 + r * c8 * pow(a8, 9) * exp(-a8 * r) / 40320;
}

/* pair type of sites i, j: 0 OO, 1 OH, 2 HH, -1 for a pair with an M site */
static int pair_type(int i, int j, int nreal)
{
    if (i >= nreal || j >= nreal) return -1;
    const int oi = (i % 3) == 0, oj = (j % 3) == 0;
    return oi && oj ? 0 : (oi || oj ? 1 : 2);
}

/* ------------------------------------------------------------------- driver ---- */

/* bond energy and dV/dr (bndfrc.f:67-108) */
static void bond_fn(int key, const double *p, double r, double *v, double *g)
{
    if (key == 1) {
        *v = 0.5 * p[0] * (r - p[1]) * (r - p[1]);
        *g = p[0] * (r - p[1]);
    } else if (key == 2) {
        const double e = exp(-p[2] * (r - p[1]));
        *v = p[0] * ((1.0 - e) * (1.0 - e) - 1.0);
        *g = 2.0 * p[0] * p[2] * (1.0 - e) * e;
    } else if (key == 7) {
        *v = p[0] * exp(-p[1] * r);
        *g = -p[0] * p[1] * exp(-p[1] * r);
    } else {
        *v = 0.0;
        *g = 0.0;
    }
}

/* force on site a due to site b (heatcurrent.f90 update_forces; W_a += F (x_a - x_b) with
 * un-imaged positions as in compute_heat_flux2) */
static void fm_add(ora_water_args *A, const double *x, const double *y, const double *z, int a, int b,
                   const double *F)
{
    if (A->force_matrix) {
        const long n = 4L * A->nw;
        for (int k = 0; k < 3; ++k) A->force_matrix[(long)k * n * n + (long)b * n + a] += F[k];
    }
    if (!A->heat_w || a == b || b >= 3 * A->nw || a >= 3 * A->nw) return;
    const double d[3] = {x[a] - x[b], y[a] - y[b], z[a] - z[b]};
    for (int p = 0; p < 3; ++p)
        for (int s = 0; s < 3; ++s) A->heat_w[9 * a + 3 * p + s] += F[p] * d[s];
}

int ora_water_terms(ora_water_args *A)
{
    if (A->imcon < 0 || A->imcon > 3) return -1;
    if (!g_ps_loaded) return -3;
    const int nw = A->nw, nreal = 3 * nw, n = 4 * nw;
    const double *x = A->x, *y = A->y, *z = A->z;
    double st[6] = {0, 0, 0, 0, 0, 0};
    /* --- monomer bonds O-H1, O-H2, H1-H2 of each water (bndfrc.f:136-330) --- */
    A->eng_bnd = 0.0;
    A->vir_bnd = 0.0;
    for (int w = 0; w < nw; ++w)
        for (int b = 0; b < 3; ++b) {
            const int ia = b == 2 ? 3 * w + 1 : 3 * w, ib = b == 2 ? 3 * w + 2 : 3 * w + 1 + b;
            const int t = b == 2 ? 1 : 0;
            if (A->bond_key[t] == 0) continue;
            double d[3] = {x[ia] - x[ib], y[ia] - y[ib], z[ia] - z[ib]};
            ora_image_vectors(A->imcon, A->cell, 1, d);
            const double rab = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            const double rrab = rab > 1e-6 ? 1.0 / rab : 0.0;
            double omega, gamma;
            bond_fn(A->bond_key[t], A->bond_prm[t], rab, &omega, &gamma);
            gamma = gamma * rrab;
            A->eng_bnd += omega;
            A->vir_bnd += gamma * rab * rab;
            if (A->atomic_energy) {
                A->atomic_energy[ia] += 0.5 * omega;
                A->atomic_energy[ib] += 0.5 * omega;
            }
            const double fv[3] = {-gamma * d[0], -gamma * d[1], -gamma * d[2]};
            const double mf[3] = {gamma * d[0], gamma * d[1], gamma * d[2]};
            for (int q = 0; q < 3; ++q) {
                A->f[3 * ia + q] += fv[q];
                A->f[3 * ib + q] -= fv[q];
            }
            fm_add(A, x, y, z, ia, ib, fv);
            fm_add(A, x, y, z, ib, ia, mf);
            st[0] += d[0] * fv[0];
            st[1] += d[0] * fv[1];
            st[2] += d[0] * fv[2];
            st[3] += d[1] * fv[1];
            st[4] += d[1] * fv[2];
            st[5] += d[2] * fv[2];
        }
    /* --- PS monomer: angle H1-O-H2 of each water (angfrc.f:160-480, keya = 9) --- */
    A->eng_ang = 0.0;
    A->vir_ang = 0.0;
    for (int w = 0; w < nw && A->ps_monomer; ++w) {
        const int ia = 3 * w + 1, ib = 3 * w, ic = 3 * w + 2;
        double dab[3] = {x[ia] - x[ib], y[ia] - y[ib], z[ia] - z[ib]};
        double dbc[3] = {x[ic] - x[ib], y[ic] - y[ib], z[ic] - z[ib]};
        ora_image_vectors(A->imcon, A->cell, 1, dab);
        ora_image_vectors(A->imcon, A->cell, 1, dbc);
        const double rab = sqrt(dab[0] * dab[0] + dab[1] * dab[1] + dab[2] * dab[2]);
        const double rbc = sqrt(dbc[0] * dbc[0] + dbc[1] * dbc[1] + dbc[2] * dbc[2]);
        const double u1[3] = {dab[0] * (1.0 / rab), dab[1] * (1.0 / rab), dab[2] * (1.0 / rab)};
        const double u2[3] = {dbc[0] * (1.0 / rbc), dbc[1] * (1.0 / rbc), dbc[2] * (1.0 / rbc)};
        const double cost = u1[0] * u2[0] + u1[1] * u2[1] + u1[2] * u2[2];
        const double theta = acos(cost);
        double pterm, vterm, fd[6];
        ora_ps_couple(rab, u1, rbc, u2, theta, &pterm, fd, &vterm);
        A->eng_ang += pterm;
        A->vir_ang += vterm;
        if (A->atomic_energy) {
            A->atomic_energy[ia] += pterm / 3;
            A->atomic_energy[ib] += pterm / 3;
            A->atomic_energy[ic] += pterm / 3;
        }
        for (int q = 0; q < 3; ++q) {
            A->f[3 * ia + q] += fd[q];
            A->f[3 * ib + q] += -fd[q] - fd[3 + q];
            A->f[3 * ic + q] += fd[3 + q];
        }
        const double mfa[3] = {-fd[0], -fd[1], -fd[2]}, mfc[3] = {-fd[3], -fd[4], -fd[5]};
        fm_add(A, x, y, z, ia, ib, fd);
        fm_add(A, x, y, z, ib, ia, mfa);
        fm_add(A, x, y, z, ic, ib, fd + 3);
        fm_add(A, x, y, z, ib, ic, mfc);
        st[0] += rab * u1[0] * fd[0] + rbc * u2[0] * fd[3];
        st[1] += rab * u1[0] * fd[1] + rbc * u2[0] * fd[4];
        st[2] += rab * u1[0] * fd[2] + rbc * u2[0] * fd[5];
        st[3] += rab * u1[1] * fd[1] + rbc * u2[1] * fd[4];
        st[4] += rab * u1[1] * fd[2] + rbc * u2[1] * fd[5];
        st[5] += rab * u1[2] * fd[2] + rbc * u2[2] * fd[5];
    }
    /* --- dispersion table (forgen.f) and pair forces (srfrce2.f) --- */
    int mxgrid = (int)(A->rcut / 0.0001 + 0.5) + 4;
    if (mxgrid < 1000) mxgrid = 1000;
    const double dlrpot = A->rvdw / (double)(mxgrid - 4);
    double *vvv = calloc((size_t)3 * (mxgrid + 3), sizeof(double)), *ggg = calloc((size_t)3 * (mxgrid + 3), sizeof(double));
    for (int t = 0; t < 3; ++t)
        for (int i = 1; i <= mxgrid; ++i) {
            const double rrr = (double)i * dlrpot;
            vvv[(size_t)t * (mxgrid + 3) + i] = vv12(rrr, A->c6[t], A->a6[t], A->c8[t], A->a8[t]);
            ggg[(size_t)t * (mxgrid + 3) + i] = gg12(rrr, A->c6[t], A->a6[t], A->c8[t], A->a8[t]);
        }
    A->eng_srp = 0.0;
    A->vir_srp = 0.0;
    const double rcsq = A->rvdw * A->rvdw, rdr = 1.0 / dlrpot;
    for (int i = 0; i < n && A->dispersion; ++i)
        for (int j = i + 1; j < n; ++j) {
            const int wi = i < nreal ? i / 3 : i - nreal, wj = j < nreal ? j / 3 : j - nreal;
            if (wi == wj) continue;
            const int t = pair_type(i, j, nreal);
            if (t < 0 || vvv[(size_t)t * (mxgrid + 3) + 1] == 0.0) continue;
            double d[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
            ora_image_vectors(A->imcon, A->cell, 1, d);
            const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (!(rcsq > rsq)) continue;
            const double rrr = sqrt(rsq);
            const int l = (int)(rrr * rdr);
            const double ppp = rrr * rdr - (double)l;
            const double *vt = vvv + (size_t)t * (mxgrid + 3), *gt = ggg + (size_t)t * (mxgrid + 3);
            double t1 = vt[l] + (vt[l + 1] - vt[l]) * ppp;
            double t2 = vt[l + 1] + (vt[l + 2] - vt[l + 1]) * (ppp - 1.0);
            const double e = t1 + (t2 - t1) * ppp * 0.5;
            A->eng_srp += e;
            if (A->atomic_energy) {
                A->atomic_energy[i] += 0.5 * e;
                A->atomic_energy[j] += 0.5 * e;
            }
            t1 = gt[l] + (gt[l + 1] - gt[l]) * ppp;
            t2 = gt[l + 1] + (gt[l + 2] - gt[l + 1]) * (ppp - 1.0);
            const double gamma = (t1 + (t2 - t1) * ppp * 0.5) / rsq;
            const double fv[3] = {gamma * d[0], gamma * d[1], gamma * d[2]};
            for (int q = 0; q < 3; ++q) {
                A->f[3 * i + q] += fv[q];
                A->f[3 * j + q] -= fv[q];
            }
            const double mf[3] = {-fv[0], -fv[1], -fv[2]};
            fm_add(A, x, y, z, j, i, mf);
            fm_add(A, x, y, z, i, j, fv);
            st[0] += d[0] * fv[0];
            st[1] += d[0] * fv[1];
            st[2] += d[0] * fv[2];
            st[3] += d[1] * fv[1];
            st[4] += d[1] * fv[2];
            st[5] += d[2] * fv[2];
            A->vir_srp = A->vir_srp - gamma * rsq;
        }
    free(vvv);
    free(ggg);
    A->stress[0] += st[0];
    A->stress[1] += st[1];
    A->stress[2] += st[2];
    A->stress[3] += st[1];
    A->stress[4] += st[3];
    A->stress[5] += st[4];
    A->stress[6] += st[2];
    A->stress[7] += st[4];
    A->stress[8] += st[5];
    /* --- long-range correction (lrcorrect.f:100-209) for types O (nw atoms) and H (2nw) --- */
    double volm;
    {
        const double *a = A->cell;
        const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[3] * (a[1] * a[8] - a[2] * a[7]) +
                           a[6] * (a[1] * a[5] - a[2] * a[4]);
        volm = fabs(det);
    }
    const double num[2] = {(double)nw, 2.0 * nw};
    const int tix[2][2] = {{0, 1}, {1, 2}};
    double elrc = 0.0, plrc = 0.0;
    /* lrcorrect.f:114: no correction without a periodic cell (imcon 0) */
    for (int ti = 0; ti < 2 && A->long_range && A->imcon != 0; ++ti)
        for (int tj = 0; tj <= ti; ++tj) {
            const int t = tix[ti][tj];
            const double r = A->rvdw;
            const double c6 = A->c6[t], a6 = A->a6[t], c8 = A->c8[t], a8 = A->a8[t];
            const double el = -c6 * pow(a6, 3) * tt_lr(6, a6 * r) - c8 * pow(a8, 5) * tt_lr(8, a8 * r);
            const double pl = 3 * el - c6 * ora_tang_toennies(6, a6 * r) / pow(r, 3) -
                              c8 * ora_tang_toennies(8, a8 * r) / pow(r, 5);
            double eadd = el, padd = pl;
            if (ti != tj) {
                eadd = eadd * 2.0;
                padd = padd * 2.0;
            }
            const double denprd = 2.0 * PI_ * (num[ti] * num[tj]) / (volm * volm);
            elrc = elrc + volm * denprd * eadd;
            plrc = plrc + denprd * padd / 3.0;
        }
    A->elrc = elrc;
    A->virlrc = plrc * (-3.0 * volm);
    return 0;
}
