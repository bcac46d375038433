/*
 * mbpol_oracle.c -- CPU restatement of mbpol_forces (the MB-pol 2B+3B driver).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restated (single rank, bead_size = 1):
 *   mbpol_forces     mbpol.f90:98-449   copy + image monomers, scan all i<j pairs and
 *                                       all i<j<k triples in lexicographic order
 *   do_dimer         mbpol.f90:453-790  predicate, re-image, leaf, E/6 per atom,
 *                                       force matrix -g/6 (all 36 entries), virial,
 *                                       forces, stress
 *   do_trimer        mbpol.f90:795-1271 predicate (>= 2 short O-O edges), re-image
 *                                       around w2, leaf, E/9, force matrix -g/9 (81)
 *   image_monomers   mbpol.f90:1275-1301
 *   image_vectors    mbpol.f90:1305-1361 (Fortran nint = round half away from zero)
 *
 * compat_inplace = 1 keeps the reference's side effect of writing the re-imaged
 * cluster back into the shared monomer array (mbpol.f90:505-509, 875-882); 0 evaluates
 * every cluster from the post-image_monomers coordinates.  The two differ only by
 * rounding of lattice shifts.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ENGUNIT 418.4 /* mbpol.f90:37 */

static ora_leaf2_fn g_leaf2 = NULL;
static ora_leaf3_fn g_leaf3 = NULL;

void ora_set_leaves(ora_leaf2_fn f2, ora_leaf3_fn f3)
{
    g_leaf2 = f2;
    g_leaf3 = f3;
}

static double leaf2(const double *w1, const double *w2, double *g1, double *g2)
{
    if (!g_leaf2) return ora_mbpol_2b(w1, w2, g1, g2);
    double e, fm[108];
    g_leaf2(w1, w2, &e, g1, g2, fm);
    return e;
}

static double leaf3(const double *w1, const double *w2, const double *w3, double *g1, double *g2, double *g3)
{
    if (!g_leaf3) return ora_mbpol_3b(w1, w2, w3, g1, g2, g3);
    double e, fm[243];
    g_leaf3(w1, w2, w3, &e, g1, g2, g3, fm);
    return e;
}

/* 3x3 inverse as DL_POLY's invert (used only for imcon 3) */
static void invert3(const double *a, double *b)
{
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
    const double r = det != 0.0 ? 1.0 / det : 0.0;
    for (int i = 0; i < 9; ++i) b[i] *= r;
}

int ora_image_vectors(int imcon, const double *cell, int n, double *dx)
{
    if (imcon == 0) return 0; /* an isolated cluster: no images */
    if (imcon == 1) {
        const double a1 = 1.0 / cell[0];
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k)
                dx[3 * i + k] = dx[3 * i + k] - cell[0] * round(dx[3 * i + k] * a1);
    } else if (imcon == 2) {
        const double r[3] = {1.0 / cell[0], 1.0 / cell[4], 1.0 / cell[8]};
        const double c[3] = {cell[0], cell[4], cell[8]};
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k)
                dx[3 * i + k] = dx[3 * i + k] - c[k] * round(dx[3 * i + k] * r[k]);
    } else if (imcon == 3) {
        double rc[9];
        invert3(cell, rc);
        for (int i = 0; i < n; ++i) {
            double *d = dx + 3 * i;
            const double ssx = rc[0] * d[0] + rc[3] * d[1] + rc[6] * d[2];
            const double ssy = rc[1] * d[0] + rc[4] * d[1] + rc[7] * d[2];
            const double ssz = rc[2] * d[0] + rc[5] * d[1] + rc[8] * d[2];
            const double xss = ssx - round(ssx), yss = ssy - round(ssy), zss = ssz - round(ssz);
            d[0] = cell[0] * xss + cell[3] * yss + cell[6] * zss;
            d[1] = cell[1] * xss + cell[4] * yss + cell[7] * zss;
            d[2] = cell[2] * xss + cell[5] * yss + cell[8] * zss;
        }
    } else {
        return -1; /* fixme stop, mbpol.f90:1357 */
    }
    return 0;
}

/* stress of one cluster: strs = -sum w (x) g with the reference's symmetrisation
 * (upper triangle copied, mbpol.f90:605-653) */
static void cluster_stress(int nmon, double *const *w, double *const *g, double *strs)
{
    double s1 = 0, s2 = 0, s3 = 0, s5 = 0, s6 = 0, s9 = 0;
    for (int m = 0; m < nmon; ++m)
        for (int a = 0; a < 3; ++a) {
            const double *p = w[m] + 3 * a, *q = g[m] + 3 * a;
            s1 += p[0] * q[0];
            s2 += p[0] * q[1];
            s3 += p[0] * q[2];
            s5 += p[1] * q[1];
            s6 += p[1] * q[2];
            s9 += p[2] * q[2];
        }
    strs[0] = -s1; strs[1] = -s2; strs[2] = -s3;
    strs[3] = -s2; strs[4] = -s5; strs[5] = -s6;
    strs[6] = -s3; strs[7] = -s6; strs[8] = -s9;
}

/* Heat-current bookkeeping of one cluster of nmon waters whose atoms have ids
 * ids[0..3*nmon-1] (1-based): force_matrix(i,j) += -g_i/k*engunit for every (i,j) in
 * the cluster (mbpol.f90:533-556, 914-943) and the per-atom contraction
 * W_i[a][b] += F_ij,a (x_i - x_j)_b with un-imaged positions (heatcurrent.f90:716-719). */
static void cluster_heat(ora_mbpol_args *A, int nmon, const int *ids, double *const *g, double kfrac)
{
    const int na = 3 * nmon;
    for (int p = 0; p < na; ++p) {
        const int i = ids[p] - 1;
        const double *gi = g[p / 3] + 3 * (p % 3);
        double f[3];
        for (int a = 0; a < 3; ++a) f[a] = -(kfrac * gi[a]) * ENGUNIT;
        for (int q = 0; q < na; ++q) {
            const int j = ids[q] - 1;
            if (A->force_matrix) {
                const long n = A->natms;
                for (int a = 0; a < 3; ++a) A->force_matrix[(long)a * n * n + (long)j * n + i] += f[a];
            }
            if (A->heat_w && j != i) {
                const double r[3] = {A->xxx[i] - A->xxx[j], A->yyy[i] - A->yyy[j], A->zzz[i] - A->zzz[j]};
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) A->heat_w[9 * i + 3 * a + b] += f[a] * r[b];
            }
        }
    }
}

static void do_dimer(ora_mbpol_args *A, double *w1, double *w2, int o1, int o2, int wi, int wj)
{
    double strs[9] = {0};
    double r12[3] = {w1[0] - w2[0], w1[1] - w2[1], w1[2] - w2[2]};
    ora_image_vectors(A->imcon, A->cell, 1, r12);
    if (r12[0] * r12[0] + r12[1] * r12[1] + r12[2] * r12[2] > 6.5 * 6.5) return;

    double c1[9], c2[9];
    double *p1 = A->compat_inplace ? w1 : c1;
    if (!A->compat_inplace) memcpy(c1, w1, sizeof c1);
    memcpy(c2, w2, sizeof c2);
    for (int k = 0; k < 3; ++k) {
        p1[3 + k] = (p1[3 + k] - p1[k]) + (w2[k] + r12[k]);
        p1[6 + k] = (p1[6 + k] - p1[k]) + (w2[k] + r12[k]);
        p1[k] = w2[k] + r12[k];
    }
    if (A->dimers && A->ndimers < A->cap_dimers) {
        A->dimers[2 * A->ndimers] = wi;
        A->dimers[2 * A->ndimers + 1] = wj;
    }
    A->ndimers++;

    double g1[9], g2[9];
    const double e2 = leaf2(p1, c2, g1, g2);
    A->eng2 += e2 * ENGUNIT;
    if (A->atomic_energy)
        for (int k = 0; k < 3; ++k) {
            A->atomic_energy[o1 - 1 + k] += (1.0 / 6.0) * e2 * ENGUNIT;
            A->atomic_energy[o2 - 1 + k] += (1.0 / 6.0) * e2 * ENGUNIT;
        }
    {
        const int ids[6] = {o1, o1 + 1, o1 + 2, o2, o2 + 1, o2 + 2};
        double *gg[2] = {g1, g2};
        cluster_heat(A, 2, ids, gg, 1.0 / 6.0);
    }
    for (int k = 0; k < 9; ++k) {
        g1[k] *= ENGUNIT;
        A->virsrp += p1[k] * g1[k];
        g2[k] *= ENGUNIT;
        A->virsrp += c2[k] * g2[k];
    }
    for (int k = 0; k < 3; ++k) {
        A->fxx[o1 - 1 + k] -= g1[3 * k];
        A->fyy[o1 - 1 + k] -= g1[3 * k + 1];
        A->fzz[o1 - 1 + k] -= g1[3 * k + 2];
        A->fxx[o2 - 1 + k] -= g2[3 * k];
        A->fyy[o2 - 1 + k] -= g2[3 * k + 1];
        A->fzz[o2 - 1 + k] -= g2[3 * k + 2];
    }
    double *ws[2] = {p1, c2}, *gs[2] = {g1, g2};
    cluster_stress(2, ws, gs, strs);
    for (int k = 0; k < 9; ++k) A->stress[k] += strs[k];
}

static void do_trimer(ora_mbpol_args *A, double *w1, double *w2, double *w3,
                      int o1, int o2, int o3, int wi, int wj, int wk)
{
    double strs[9] = {0};
    double rab[9];
    for (int k = 0; k < 3; ++k) {
        rab[k] = w1[k] - w2[k];
        rab[k + 3] = w2[k] - w3[k];
        rab[k + 6] = w1[k] - w3[k];
    }
    ora_image_vectors(A->imcon, A->cell, 3, rab);
    double r12 = 0, r23 = 0, r13 = 0;
    for (int k = 0; k < 3; ++k) {
        r12 += rab[k] * rab[k];
        r23 += rab[k + 3] * rab[k + 3];
        r13 += rab[k + 6] * rab[k + 6];
    }
    r12 = sqrt(r12);
    r23 = sqrt(r23);
    r13 = sqrt(r13);
    const double r3f = 4.5;
    const int big1 = r12 > r3f || r13 > r3f;
    const int big2 = r12 > r3f || r23 > r3f;
    const int big3 = r13 > r3f || r23 > r3f;
    if (big1 && big2 && big3) return;

    double c1[9], c2[9], c3[9];
    double *p1 = A->compat_inplace ? w1 : c1;
    double *p3 = A->compat_inplace ? w3 : c3;
    if (!A->compat_inplace) {
        memcpy(c1, w1, sizeof c1);
        memcpy(c3, w3, sizeof c3);
    }
    memcpy(c2, w2, sizeof c2);
    for (int k = 0; k < 3; ++k) {
        p1[3 + k] = (p1[3 + k] - p1[k]) + (w2[k] + rab[k]);
        p1[6 + k] = (p1[6 + k] - p1[k]) + (w2[k] + rab[k]);
        p1[k] = w2[k] + rab[k];
        p3[3 + k] = (p3[3 + k] - p3[k]) + (w2[k] - rab[k + 3]);
        p3[6 + k] = (p3[6 + k] - p3[k]) + (w2[k] - rab[k + 3]);
        p3[k] = w2[k] - rab[k + 3];
    }
    if (A->trimers && A->ntrimers < A->cap_trimers) {
        A->trimers[3 * A->ntrimers] = wi;
        A->trimers[3 * A->ntrimers + 1] = wj;
        A->trimers[3 * A->ntrimers + 2] = wk;
    }
    A->ntrimers++;

    double g1[9] = {0}, g2[9] = {0}, g3[9] = {0};
    const double e3 = leaf3(p1, c2, p3, g1, g2, g3);
    A->eng3 += e3 * ENGUNIT;
    if (A->atomic_energy)
        for (int k = 0; k < 3; ++k) {
            A->atomic_energy[o1 - 1 + k] += (1.0 / 9.0) * e3 * ENGUNIT;
            A->atomic_energy[o2 - 1 + k] += (1.0 / 9.0) * e3 * ENGUNIT;
            A->atomic_energy[o3 - 1 + k] += (1.0 / 9.0) * e3 * ENGUNIT;
        }
    {
        const int ids[9] = {o1, o1 + 1, o1 + 2, o2, o2 + 1, o2 + 2, o3, o3 + 1, o3 + 2};
        double *gg[3] = {g1, g2, g3};
        cluster_heat(A, 3, ids, gg, 1.0 / 9.0);
    }
    for (int k = 0; k < 9; ++k) {
        g1[k] *= ENGUNIT;
        A->virsrp += p1[k] * g1[k];
        g2[k] *= ENGUNIT;
        A->virsrp += c2[k] * g2[k];
        g3[k] *= ENGUNIT;
        A->virsrp += p3[k] * g3[k];
    }
    for (int k = 0; k < 3; ++k) {
        A->fxx[o1 - 1 + k] -= g1[3 * k];
        A->fyy[o1 - 1 + k] -= g1[3 * k + 1];
        A->fzz[o1 - 1 + k] -= g1[3 * k + 2];
        A->fxx[o2 - 1 + k] -= g2[3 * k];
        A->fyy[o2 - 1 + k] -= g2[3 * k + 1];
        A->fzz[o2 - 1 + k] -= g2[3 * k + 2];
        A->fxx[o3 - 1 + k] -= g3[3 * k];
        A->fyy[o3 - 1 + k] -= g3[3 * k + 1];
        A->fzz[o3 - 1 + k] -= g3[3 * k + 2];
    }
    double *ws[3] = {p1, c2, p3}, *gs[3] = {g1, g2, g3};
    cluster_stress(3, ws, gs, strs);
    for (int k = 0; k < 9; ++k) A->stress[k] += strs[k];
}

int ora_mbpol_forces(ora_mbpol_args *A)
{
    if (A->imcon < 0 || A->imcon > 3) return -1;
    const int nw = A->nw;
    double *mon = malloc(sizeof(double) * 9 * (nw > 0 ? nw : 1));
    for (int t = 0; t < 3 * nw; ++t) {
        const int i = A->listttm2[t] - 1;
        mon[3 * t] = A->xxx[i];
        mon[3 * t + 1] = A->yyy[i];
        mon[3 * t + 2] = A->zzz[i];
    }
    for (int i = 0; i < nw; ++i) {
        double oh[6];
        double *m = mon + 9 * i;
        for (int k = 0; k < 3; ++k) {
            oh[k] = m[3 + k] - m[k];
            oh[k + 3] = m[6 + k] - m[k];
        }
        ora_image_vectors(A->imcon, A->cell, 2, oh);
        for (int k = 0; k < 3; ++k) {
            m[3 + k] = m[k] + oh[k];
            m[6 + k] = m[k] + oh[k + 3];
        }
    }
    A->eng2 = A->eng3 = 0.0;
    A->ndimers = A->ntrimers = 0;
    double vir0 = A->virsrp;
    for (int i = 0; i < nw; ++i)
        for (int j = i + 1; j < nw; ++j)
            do_dimer(A, mon + 9 * i, mon + 9 * j, A->listttm2[3 * i], A->listttm2[3 * j], i, j);
    if (!A->skip_3b)
        for (int i = 0; i < nw; ++i)
            for (int j = i + 1; j < nw; ++j)
                for (int k = j + 1; k < nw; ++k)
                    do_trimer(A, mon + 9 * i, mon + 9 * j, mon + 9 * k, A->listttm2[3 * i],
                              A->listttm2[3 * j], A->listttm2[3 * k], i, j, k);
    (void)vir0;
    A->engsrp += A->eng2 + A->eng3;
    free(mon);
    return 0;
}
