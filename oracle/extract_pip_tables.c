/*
 * extract_pip_tables.c -- recover the monomial content of the reference's MB-pol
 * permutationally invariant polynomials by black-box probing of the compiled
 * reference leaves (oracle/_ref/libmbpol_ref.so).
 *
 * Test/data infrastructure, run once in a container that holds the reference:
 * its output (a monomial table per polynomial) is committed as data and the
 * product never runs this program.
 *
 * What is probed.  poly_2b_v6x::eval (poly-2b-v6x.h:10-18) and poly_3b_v2x::eval
 * (poly-3b-v2x.h:10-17) compute E(a, x) = sum_i a_i p_i(x) together with dE/dx,
 * where every basis polynomial p_i has degree <= 4 (poly-2b-v6x.h:11) and the
 * generated code uses only +, * and small integer literals.  With a = e_i and
 * small integer x the double arithmetic is exact, so p_i can be interpolated
 * exactly:
 *   1. the variables p_i depends on come from the gradient at a random point;
 *   2. monomials are found by exact-support sets S of growing size: the value
 *      at integer points supported on S minus the already-known monomials with
 *      smaller support leaves only the monomials whose support is exactly S,
 *      solved from a small linear system (|S| = 1..3);
 *   3. |S| = 4 (x_a x_b x_c x_d) comes from the gradient at the indicator point
 *      of the three lower variables, minus the known lower-support terms.
 * The result is checked afterwards against eval() with random a and random real
 * x (value and all partials), which is what pins the table.
 *
 * Output (stdout): one line per monomial: "basis coef v1 [v2 [v3 [v4]]]" with
 * the variable indices repeated by exponent, ascending.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

double _ZN3x2o11poly_2b_v6x4evalEPKdS2_Pd(const double *, const double *, double *);
double _ZN3x2o11poly_3b_v2x4evalEPKdS2_Pd(const double *, const double *, double *);

static int g_which, g_nv, g_na;

static double peval(const double *a, const double *x, double *g)
{
    return g_which == 2 ? _ZN3x2o11poly_2b_v6x4evalEPKdS2_Pd(a, x, g)
                        : _ZN3x2o11poly_3b_v2x4evalEPKdS2_Pd(a, x, g);
}

typedef struct {
    int coef;
    int ex[36];
} mono_t;

typedef struct {
    mono_t *m;
    int n, cap;
} mlist_t;

static void push(mlist_t *L, const int *ex, int coef)
{
    if (L->n == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 64;
        L->m = realloc(L->m, sizeof(mono_t) * L->cap);
    }
    L->m[L->n].coef = coef;
    memcpy(L->m[L->n].ex, ex, sizeof(int) * 36);
    L->n++;
}

static double mono_val(const mono_t *m, const double *x)
{
    double v = m->coef;
    for (int j = 0; j < g_nv; ++j)
        for (int e = 0; e < m->ex[j]; ++e) v *= x[j];
    return v;
}

/* value of the known monomials at x */
static double known_val(const mlist_t *L, const double *x)
{
    double s = 0;
    for (int t = 0; t < L->n; ++t) s += mono_val(&L->m[t], x);
    return s;
}

/* Solve the small dense system A c = b (n <= 6) by Gaussian elimination with
 * partial pivoting; the entries are exact integers, the solution is rounded
 * and checked for integrality by the caller. */
static int solve(int n, double A[][6], double *b, double *c)
{
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int r = k + 1; r < n; ++r)
            if (fabs(A[r][k]) > fabs(A[p][k])) p = r;
        if (fabs(A[p][k]) < 1e-12) return -1;
        for (int q = 0; q < n; ++q) { double t = A[k][q]; A[k][q] = A[p][q]; A[p][q] = t; }
        { double t = b[k]; b[k] = b[p]; b[p] = t; }
        for (int r = k + 1; r < n; ++r) {
            double f = A[r][k] / A[k][k];
            for (int q = k; q < n; ++q) A[r][q] -= f * A[k][q];
            b[r] -= f * b[k];
        }
    }
    for (int k = n - 1; k >= 0; --k) {
        double s = b[k];
        for (int q = k + 1; q < n; ++q) s -= A[k][q] * c[q];
        c[k] = s / A[k][k];
    }
    return 0;
}

/* exponent patterns for an exact support of size k (all exponents >= 1, degree <= 4) */
static const int PAT1[4][1] = {{1}, {2}, {3}, {4}};
static const int PAT2[6][2] = {{1, 1}, {1, 2}, {2, 1}, {1, 3}, {3, 1}, {2, 2}};
static const int PAT3[4][3] = {{1, 1, 1}, {2, 1, 1}, {1, 2, 1}, {1, 1, 2}};
static const int PTS1[4][1] = {{1}, {2}, {3}, {4}};
static const int PTS2[6][2] = {{1, 1}, {1, 2}, {2, 1}, {2, 2}, {1, 3}, {3, 1}};
static const int PTS3[4][3] = {{1, 1, 1}, {2, 1, 1}, {1, 2, 1}, {1, 1, 2}};

static int fail = 0;

static void exact_support(int i, mlist_t *L, int k, const int *S)
{
    int np = k == 1 ? 4 : (k == 2 ? 6 : 4);
    double A[6][6], b[6], c[6];
    double a[1163], x[36], g[36];
    memset(a, 0, sizeof a);
    a[i] = 1.0;
    for (int p = 0; p < np; ++p) {
        memset(x, 0, sizeof x);
        for (int q = 0; q < k; ++q)
            x[S[q]] = k == 1 ? PTS1[p][q] : (k == 2 ? PTS2[p][q] : PTS3[p][q]);
        double v = peval(a, x, g);
        b[p] = v - known_val(L, x);
        for (int u = 0; u < np; ++u) {
            double m = 1.0;
            for (int q = 0; q < k; ++q) {
                int e = k == 1 ? PAT1[u][q] : (k == 2 ? PAT2[u][q] : PAT3[u][q]);
                for (int r = 0; r < e; ++r) m *= x[S[q]];
            }
            A[p][u] = m;
        }
    }
    int allzero = 1;
    for (int p = 0; p < np; ++p) if (b[p] != 0.0) allzero = 0;
    if (allzero) return;
    if (solve(np, A, b, c)) { fprintf(stderr, "singular system\n"); fail = 1; return; }
    int ex[36];
    for (int u = 0; u < np; ++u) {
        long ci = lround(c[u]);
        if (fabs(c[u] - ci) > 1e-6) {
            fprintf(stderr, "non-integer coefficient %g (basis %d)\n", c[u], i);
            fail = 1;
        }
        if (ci == 0) continue;
        memset(ex, 0, sizeof ex);
        for (int q = 0; q < k; ++q)
            ex[S[q]] = k == 1 ? PAT1[u][q] : (k == 2 ? PAT2[u][q] : PAT3[u][q]);
        push(L, ex, (int)ci);
    }
}

static void basis_monomials(int i, mlist_t *L)
{
    double a[1163], x[36], g[36], gs[36];
    int V[36], nv = 0;
    memset(a, 0, sizeof a);
    a[i] = 1.0;
    unsigned short seed[3] = {(unsigned short)(i + 1), 7, 11};
    memset(gs, 0, sizeof gs);
    for (int rep = 0; rep < 2; ++rep) {
        for (int j = 0; j < g_nv; ++j) x[j] = 0.3 + 1.3 * erand48(seed);
        peval(a, x, g);
        for (int j = 0; j < g_nv; ++j) gs[j] += fabs(g[j]);
    }
    for (int j = 0; j < g_nv; ++j)
        if (gs[j] != 0.0) V[nv++] = j;

    /* constant term */
    memset(x, 0, sizeof x);
    double c0 = peval(a, x, g);
    if (c0 != 0.0) {
        int ex[36] = {0};
        push(L, ex, (int)lround(c0));
    }
    int S[4];
    for (int p = 0; p < nv; ++p) { S[0] = V[p]; exact_support(i, L, 1, S); }
    for (int p = 0; p < nv; ++p)
        for (int q = p + 1; q < nv; ++q) { S[0] = V[p]; S[1] = V[q]; exact_support(i, L, 2, S); }
    for (int p = 0; p < nv; ++p)
        for (int q = p + 1; q < nv; ++q)
            for (int r = q + 1; r < nv; ++r) {
                S[0] = V[p]; S[1] = V[q]; S[2] = V[r];
                exact_support(i, L, 3, S);
            }
    /* quadrilinear terms x_a x_b x_c x_m with m > c, from the gradient at 1_{a,b,c} */
    int nlow = L->n;
    for (int p = 0; p < nv; ++p)
        for (int q = p + 1; q < nv; ++q)
            for (int r = q + 1; r < nv; ++r) {
                memset(x, 0, sizeof x);
                x[V[p]] = x[V[q]] = x[V[r]] = 1.0;
                peval(a, x, g);
                for (int s = r + 1; s < nv; ++s) {
                    int m = V[s];
                    /* known part of dp/dx_m at this point: monomials with ex[m]==1 and
                     * support inside {a,b,c,m}, all of support size <= 3 */
                    double kn = 0.0;
                    for (int t = 0; t < nlow; ++t) {
                        const mono_t *mm = &L->m[t];
                        if (mm->ex[m] != 1) continue;
                        int ok = 1;
                        for (int j = 0; j < g_nv; ++j)
                            if (mm->ex[j] && j != m && x[j] == 0.0) { ok = 0; break; }
                        if (ok) kn += mm->coef;
                    }
                    double d = g[m] - kn;
                    long ci = lround(d);
                    if (fabs(d - ci) > 1e-6) {
                        fprintf(stderr, "non-integer quad coefficient %g (basis %d)\n", d, i);
                        fail = 1;
                    }
                    if (ci != 0) {
                        int ex[36] = {0};
                        ex[V[p]] = ex[V[q]] = ex[V[r]] = ex[m] = 1;
                        push(L, ex, (int)ci);
                    }
                }
            }
}

int main(int argc, char **argv)
{
    if (argc < 2) { fprintf(stderr, "usage: %s 2|3\n", argv[0]); return 2; }
    g_which = atoi(argv[1]);
    g_nv = g_which == 2 ? 31 : 36;
    g_na = g_which == 2 ? 1153 : 1163;
    mlist_t *all = calloc(g_na, sizeof(mlist_t));
    srand48(12345);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < g_na; ++i) basis_monomials(i, &all[i]);
    if (fail) { fprintf(stderr, "extraction failed\n"); return 1; }

    /* verification against eval() at random a and random real x */
    double maxrel = 0.0;
    for (int rep = 0; rep < 20; ++rep) {
        double a[1163], x[36], g[36], gr[36];
        for (int i = 0; i < g_na; ++i) a[i] = 2.0 * drand48() - 1.0;
        for (int j = 0; j < g_nv; ++j) x[j] = 0.05 + 1.2 * drand48();
        double e = peval(a, x, g);
        double er = 0.0, scale = 0.0;
        memset(gr, 0, sizeof gr);
        for (int i = 0; i < g_na; ++i)
            for (int t = 0; t < all[i].n; ++t) {
                const mono_t *m = &all[i].m[t];
                double v = a[i] * mono_val(m, x);
                er += v;
                scale += fabs(v);
                for (int j = 0; j < g_nv; ++j)
                    if (m->ex[j]) gr[j] += v * m->ex[j] / x[j];
            }
        double rel = fabs(e - er) / scale;
        if (rel > maxrel) maxrel = rel;
        for (int j = 0; j < g_nv; ++j) {
            double r2 = fabs(g[j] - gr[j]) / (scale / x[j]);
            if (r2 > maxrel) maxrel = r2;
        }
    }
    long nm = 0;
    for (int i = 0; i < g_na; ++i) nm += all[i].n;
    fprintf(stderr, "poly %db: %d basis polynomials, %ld monomials, max rel. deviation %.3e\n",
            g_which, g_na, nm, maxrel);
    if (maxrel > 1e-13) { fprintf(stderr, "verification failed\n"); return 1; }

    for (int i = 0; i < g_na; ++i)
        for (int t = 0; t < all[i].n; ++t) {
            const mono_t *m = &all[i].m[t];
            printf("%d %d", i, m->coef);
            for (int j = 0; j < g_nv; ++j)
                for (int e = 0; e < m->ex[j]; ++e) printf(" %d", j);
            printf("\n");
        }
    return 0;
}
