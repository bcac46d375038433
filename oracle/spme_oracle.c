/*
 * spme_oracle.c -- CPU restatement of the smooth particle-mesh Ewald reciprocal space for
 * point charges and point dipoles, and the plain Ewald reciprocal sum it approximates.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restated (reference file:line):
 *   cardinal B-splines M_n(w + t) by the recursion in n        bspgen.f:27-100
 *   |b(m)|^2 moduli of the interpolation                        bspcoe.f:1-95
 *   charge spreading, FFT, influence function, gather           ewald_spme.f:1-362, spme_for.f
 * The reference's SPME spreads charges only; here the spread density also carries the
 * dipoles (mu . grad of the spline product), as the induced-dipole Ewald sum of the path
 * (ewald1p.f) needs.  Its reciprocal sum over k is restated directly in ora_ewald_recip and
 * is what SPME approximates.
 *
 * Units: Coulomb constant 1 (energy e^2/A, potential e/A, field e/A^2, force e^2/A^2);
 * the caller multiplies by r4pie0/epsq.  The cell is three rows (cell[0..2] = first
 * vector, as DL_POLY's cell(1..3)); fractional coordinates s_a = a*_a . r with a*_a the
 * reciprocal vectors (a*_a . a_b = delta_ab).
 *
 * Per site i the outputs are phi_i (potential), E_i = -grad phi_i (field) and the force
 * F_i = q_i E_i + (mu_i . grad) E_i = -dU/dr_i, every one including the site's own image
 * sum, as a reciprocal-space sum does; U = 1/2 sum_i (q_i phi_i - mu_i . E_i).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define PI_ 3.141592653589793

static void recip_vectors(const double *cell, double rv[3][3], double *vol)
{
    const double *a = cell, *b = cell + 3, *c = cell + 6;
    double bxc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    double cxa[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
    double axb[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double v = a[0] * bxc[0] + a[1] * bxc[1] + a[2] * bxc[2];
    for (int d = 0; d < 3; ++d) {
        rv[0][d] = bxc[d] / v;
        rv[1][d] = cxa[d] / v;
        rv[2][d] = axb[d] / v;
    }
    *vol = fabs(v);
}

/* a[t] = M_p(w + t), d[t] = M_p'(w + t), d2[t] = M_p''(w + t) for t = 0..p-1, 0 <= w < 1
 * (bspgen.f:27-100: the same recursion M_n(x) = [x M_{n-1}(x) + (n - x) M_{n-1}(x-1)]/(n-1)
 * from M_2, with M_n' = M_{n-1}(x) - M_{n-1}(x-1)). */
void ora_bspline(int p, double w, double *a, double *d, double *d2)
{
    double m[32];
    double m1[32], m2[32];   /* orders p-1 and p-2 */
    memset(m, 0, sizeof m);
    m[0] = w;
    m[1] = 1.0 - w;
    for (int n = 3; n <= p; ++n) {
        if (n == p - 1) memcpy(m2, m, sizeof m);
        if (n == p) memcpy(m1, m, sizeof m);
        double nx[32];
        memset(nx, 0, sizeof nx);
        for (int t = 0; t < n; ++t) {
            const double x = w + t;
            const double cur = t < n - 1 ? m[t] : 0.0;
            const double prev = t > 0 ? m[t - 1] : 0.0;
            nx[t] = (x * cur + ((double)n - x) * prev) / (double)(n - 1);
        }
        memcpy(m, nx, sizeof m);
    }
    for (int t = 0; t < p; ++t) {
        a[t] = m[t];
        d[t] = (t < p - 1 ? m1[t] : 0.0) - (t > 0 ? m1[t - 1] : 0.0);
        d2[t] = (t < p - 2 ? m2[t] : 0.0) - 2.0 * (t > 0 && t - 1 < p - 2 ? m2[t - 1] : 0.0) +
                (t > 1 ? m2[t - 2] : 0.0);
    }
}

/* |b(m)|^2 = 1 / |sum_{k=0}^{p-2} M_p(k+1) exp(2 pi i m k / K)|^2 (bspcoe.f) */
void ora_spme_bmod(int p, int K, double *bmod)
{
    double a[32], d[32], d2[32];
    ora_bspline(p, 0.0, a, d, d2);   /* a[t] = M_p(t) */
    for (int m = 0; m < K; ++m) {
        double re = 0.0, im = 0.0;
        for (int k = 0; k <= p - 2; ++k) {
            const double ang = 2.0 * PI_ * (double)m * (double)k / (double)K;
            re += a[k + 1] * cos(ang);
            im += a[k + 1] * sin(ang);
        }
        bmod[m] = 1.0 / (re * re + im * im);
    }
}

/* one-dimensional DFT along an axis of a complex grid [K0][K1][K2] (interleaved re, im),
 * sign -1 forward, +1 backward, unnormalised; O(K^2) per line */
static void dft_axis(double *g, const int *K, int axis, int sign)
{
    const int n = K[axis];
    const long st[3] = {(long)K[1] * K[2], K[2], 1};
    const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
    double *line = malloc(sizeof(double) * 2 * n), *out = malloc(sizeof(double) * 2 * n);
    for (int i = 0; i < K[o1]; ++i)
        for (int j = 0; j < K[o2]; ++j) {
            const long base = i * st[o1] + j * st[o2];
            for (int e = 0; e < n; ++e) {
                line[2 * e] = g[2 * (base + e * st[axis])];
                line[2 * e + 1] = g[2 * (base + e * st[axis]) + 1];
            }
            for (int m = 0; m < n; ++m) {
                double re = 0.0, im = 0.0;
                for (int e = 0; e < n; ++e) {
                    const double ang = sign * 2.0 * PI_ * (double)(((long)m * e) % n) / (double)n;
                    const double c = cos(ang), s = sin(ang);
                    re += line[2 * e] * c - line[2 * e + 1] * s;
                    im += line[2 * e] * s + line[2 * e + 1] * c;
                }
                out[2 * m] = re;
                out[2 * m + 1] = im;
            }
            for (int e = 0; e < n; ++e) {
                g[2 * (base + e * st[axis])] = out[2 * e];
                g[2 * (base + e * st[axis]) + 1] = out[2 * e + 1];
            }
        }
    free(line);
    free(out);
}

static inline int wrap(int g, int K) { return ((g % K) + K) % K; }

/* SPME reciprocal space for charges q[n] and dipoles mu[3n] (NULL: none) at positions
 * x, y, z; order p (even, 4..12), grid K[3].  Outputs (each may be NULL): phi[n], E[3n],
 * F[3n]; returns U. */
double ora_spme_recip(const double *cell, int p, const int *K, double alpha, int n, const double *x,
                      const double *y, const double *z, const double *q, const double *mu, double *phi, double *E,
                      double *F)
{
    double rv[3][3], vol;
    recip_vectors(cell, rv, &vol);
    const long ng = (long)K[0] * K[1] * K[2];
    double *Q = calloc(2 * ng, sizeof(double));
    double *ba[3];
    for (int a = 0; a < 3; ++a) {
        ba[a] = malloc(sizeof(double) * K[a]);
        ora_spme_bmod(p, K[a], ba[a]);
    }
    /* per-site splines */
    double *th = malloc(sizeof(double) * (size_t)n * 9 * p);
    int *g0 = malloc(sizeof(int) * 3 * n);
    for (int i = 0; i < n; ++i) {
        const double r[3] = {x[i], y[i], z[i]};
        for (int a = 0; a < 3; ++a) {
            const double u = K[a] * (rv[a][0] * r[0] + rv[a][1] * r[1] + rv[a][2] * r[2]);
            const double fl = floor(u);
            g0[3 * i + a] = (int)fl;
            double *t = th + ((size_t)i * 9 + 3 * a) * p;
            ora_bspline(p, u - fl, t, t + p, t + 2 * p);
        }
    }
    /* spread: Q(g) += q theta1 theta2 theta3 + mu . grad(theta1 theta2 theta3), grid
     * point g_a = floor(u_a) - t_a for weight M_p(w_a + t_a) */
    for (int i = 0; i < n; ++i) {
        double wv[3] = {0, 0, 0};
        if (mu)
            for (int a = 0; a < 3; ++a)
                wv[a] = K[a] * (rv[a][0] * mu[3 * i] + rv[a][1] * mu[3 * i + 1] + rv[a][2] * mu[3 * i + 2]);
        const double *t1 = th + (size_t)i * 9 * p, *t2 = t1 + 3 * p, *t3 = t1 + 6 * p;
        for (int a1 = 0; a1 < p; ++a1)
            for (int a2 = 0; a2 < p; ++a2)
                for (int a3 = 0; a3 < p; ++a3) {
                    const double c = q[i] * t1[a1] * t2[a2] * t3[a3] + wv[0] * t1[p + a1] * t2[a2] * t3[a3] +
                                     wv[1] * t1[a1] * t2[p + a2] * t3[a3] + wv[2] * t1[a1] * t2[a2] * t3[p + a3];
                    const long gi = ((long)wrap(g0[3 * i] - a1, K[0]) * K[1] + wrap(g0[3 * i + 1] - a2, K[1])) * K[2] +
                                    wrap(g0[3 * i + 2] - a3, K[2]);
                    Q[2 * gi] += c;
                }
    }
    /* forward transform, influence function, energy, backward transform */
    for (int a = 0; a < 3; ++a) dft_axis(Q, K, a, -1);
    double U = 0.0;
    for (int m1 = 0; m1 < K[0]; ++m1)
        for (int m2 = 0; m2 < K[1]; ++m2)
            for (int m3 = 0; m3 < K[2]; ++m3) {
                const long gi = ((long)m1 * K[1] + m2) * K[2] + m3;
                const int mm[3] = {m1 <= K[0] / 2 ? m1 : m1 - K[0], m2 <= K[1] / 2 ? m2 : m2 - K[1],
                                   m3 <= K[2] / 2 ? m3 : m3 - K[2]};
                double bc = 0.0;
                if (mm[0] || mm[1] || mm[2]) {
                    double mv[3];
                    for (int d = 0; d < 3; ++d) mv[d] = mm[0] * rv[0][d] + mm[1] * rv[1][d] + mm[2] * rv[2][d];
                    const double m2v = mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2];
                    bc = ba[0][m1] * ba[1][m2] * ba[2][m3] * exp(-PI_ * PI_ * m2v / (alpha * alpha)) / (PI_ * vol * m2v);
                }
                U += 0.5 * bc * (Q[2 * gi] * Q[2 * gi] + Q[2 * gi + 1] * Q[2 * gi + 1]);
                Q[2 * gi] *= bc;
                Q[2 * gi + 1] *= bc;
            }
    for (int a = 0; a < 3; ++a) dft_axis(Q, K, a, +1);
    /* gather phi, grad phi and the Hessian of phi at every site */
    for (int i = 0; i < n; ++i) {
        const double *t1 = th + (size_t)i * 9 * p, *t2 = t1 + 3 * p, *t3 = t1 + 6 * p;
        double ph = 0.0, gu[3] = {0, 0, 0}, hu[3][3] = {{0}};
        for (int a1 = 0; a1 < p; ++a1)
            for (int a2 = 0; a2 < p; ++a2)
                for (int a3 = 0; a3 < p; ++a3) {
                    const long gi = ((long)wrap(g0[3 * i] - a1, K[0]) * K[1] + wrap(g0[3 * i + 1] - a2, K[1])) * K[2] +
                                    wrap(g0[3 * i + 2] - a3, K[2]);
                    const double v = Q[2 * gi];
                    const double A[3] = {t1[a1], t2[a2], t3[a3]}, D[3] = {t1[p + a1], t2[p + a2], t3[p + a3]},
                                 D2[3] = {t1[2 * p + a1], t2[2 * p + a2], t3[2 * p + a3]};
                    ph += v * A[0] * A[1] * A[2];
                    gu[0] += v * D[0] * A[1] * A[2];
                    gu[1] += v * A[0] * D[1] * A[2];
                    gu[2] += v * A[0] * A[1] * D[2];
                    hu[0][0] += v * D2[0] * A[1] * A[2];
                    hu[1][1] += v * A[0] * D2[1] * A[2];
                    hu[2][2] += v * A[0] * A[1] * D2[2];
                    hu[0][1] += v * D[0] * D[1] * A[2];
                    hu[0][2] += v * D[0] * A[1] * D[2];
                    hu[1][2] += v * A[0] * D[1] * D[2];
                }
        hu[1][0] = hu[0][1];
        hu[2][0] = hu[0][2];
        hu[2][1] = hu[1][2];
        /* d/dr_d = sum_a K_a a*_a[d] d/du_a */
        double g[3], H[3][3];
        for (int d = 0; d < 3; ++d) {
            g[d] = 0.0;
            for (int a = 0; a < 3; ++a) g[d] += K[a] * rv[a][d] * gu[a];
        }
        for (int d = 0; d < 3; ++d)
            for (int e = 0; e < 3; ++e) {
                H[d][e] = 0.0;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) H[d][e] += K[a] * rv[a][d] * K[b] * rv[b][e] * hu[a][b];
            }
        if (phi) phi[i] = ph;
        if (E)
            for (int d = 0; d < 3; ++d) E[3 * i + d] = -g[d];
        if (F)
            for (int d = 0; d < 3; ++d) {
                double hm = 0.0;
                if (mu)
                    for (int e = 0; e < 3; ++e) hm += H[d][e] * mu[3 * i + e];
                F[3 * i + d] = -q[i] * g[d] - hm;
            }
    }
    free(Q);
    free(th);
    free(g0);
    for (int a = 0; a < 3; ++a) free(ba[a]);
    return U;
}

/* The Ewald reciprocal sum the above approximates (ewald1p.f's reciprocal part with
 * S(k) = sum_j (q_j + i k . mu_j) exp(i k . r_j)): U = (2 pi / V) sum_{k != 0, |m_a| <=
 * kmax_a} exp(-k^2 / 4 alpha^2) / k^2 |S(k)|^2 with k = 2 pi sum_a m_a a*_a; phi, E and F
 * as for ora_spme_recip. */
double ora_ewald_recip(const double *cell, const int *kmax, double alpha, int n, const double *x, const double *y,
                       const double *z, const double *q, const double *mu, double *phi, double *E, double *F)
{
    double rv[3][3], vol;
    recip_vectors(cell, rv, &vol);
    double *cs = malloc(sizeof(double) * 2 * n);
    if (phi) memset(phi, 0, sizeof(double) * n);
    if (E) memset(E, 0, sizeof(double) * 3 * n);
    if (F) memset(F, 0, sizeof(double) * 3 * n);
    double U = 0.0;
    for (int m1 = -kmax[0]; m1 <= kmax[0]; ++m1)
        for (int m2 = -kmax[1]; m2 <= kmax[1]; ++m2)
            for (int m3 = -kmax[2]; m3 <= kmax[2]; ++m3) {
                if (!m1 && !m2 && !m3) continue;
                double k[3];
                for (int d = 0; d < 3; ++d) k[d] = 2.0 * PI_ * (m1 * rv[0][d] + m2 * rv[1][d] + m3 * rv[2][d]);
                const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                const double f = 4.0 * PI_ / vol * exp(-k2 / (4.0 * alpha * alpha)) / k2;
                double sr = 0.0, si = 0.0;
                for (int j = 0; j < n; ++j) {
                    const double kr = k[0] * x[j] + k[1] * y[j] + k[2] * z[j];
                    const double c = cos(kr), s = sin(kr);
                    const double kmu = mu ? k[0] * mu[3 * j] + k[1] * mu[3 * j + 1] + k[2] * mu[3 * j + 2] : 0.0;
                    /* (q + i kmu)(c + i s) */
                    sr += q[j] * c - kmu * s;
                    si += q[j] * s + kmu * c;
                    cs[2 * j] = c;
                    cs[2 * j + 1] = s;
                }
                U += 0.5 * f * (sr * sr + si * si);
                for (int i = 0; i < n; ++i) {
                    /* Re[S e^{-ik.r}] and Im[S e^{-ik.r}] */
                    const double c = cs[2 * i], s = cs[2 * i + 1];
                    const double re = sr * c + si * s, im = si * c - sr * s;
                    if (phi) phi[i] += f * re;
                    /* E = -grad phi = f Re[i k S e^{-ik.r}] = -f k Im[...] */
                    if (E)
                        for (int d = 0; d < 3; ++d) E[3 * i + d] += -f * k[d] * im;
                    if (F) {
                        /* grad grad phi = -f k k Re[...]; F = -q grad phi - H mu */
                        const double kmu = mu ? k[0] * mu[3 * i] + k[1] * mu[3 * i + 1] + k[2] * mu[3 * i + 2] : 0.0;
                        for (int d = 0; d < 3; ++d) F[3 * i + d] += -q[i] * f * k[d] * im + f * k[d] * kmu * re;
                    }
                }
            }
    free(cs);
    return U;
}
