/*
 * heat_oracle.c -- CPU restatement of the Green-Kubo heat flux (compute_heat_flux2).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   compute_heat_flux2   heatcurrent.f90:703-728
 *     J_kin = sum_i (e_kin,i + e_pot,i) v_i
 *     J_pot = sum_i [ sum_{j != i} (x_i - x_j) (F_ij . v_i) - sigma^{ew1p}_i . v_i ]
 *   over the real atoms only (the first 3/4 of the site arrays), with un-imaged
 *   positions.  The dense version reads the full force matrix; the tensor version reads
 *   the per-atom contraction W_i[a][b] = sum_j F_ij,a (x_i - x_j)_b, which is the same
 *   sum regrouped: J_pot,b = sum_i sum_a v_i,a W_i[a][b].
 */
#include <string.h>

#include "oracle.h"

static void ew1p_term(const double *s, const double *v, double *jp)
{
    /* the reference uses only the upper triangle of sigma (heatcurrent.f90:720-725) */
    jp[0] -= s[0] * v[0] + s[1] * v[1] + s[2] * v[2];
    jp[1] -= s[1] * v[0] + s[4] * v[1] + s[5] * v[2];
    jp[2] -= s[2] * v[0] + s[5] * v[1] + s[8] * v[2];
}

void ora_heat_flux_dense(int natms, int nreal, const double *fm,
                         const double *x, const double *y, const double *z,
                         const double *vx, const double *vy, const double *vz,
                         const double *e_kin, const double *e_pot,
                         const double *stress_ew1p,
                         double jkin[3], double jpot[3], double jtot[3])
{
    const long n = natms;
    memset(jkin, 0, 3 * sizeof(double));
    memset(jpot, 0, 3 * sizeof(double));
    for (int i = 0; i < nreal; ++i) {
        const double v[3] = {vx[i], vy[i], vz[i]};
        const double eng = e_kin[i] + e_pot[i];
        for (int a = 0; a < 3; ++a) jkin[a] += eng * v[a];
        for (int j = 0; j < nreal; ++j) {
            if (i == j) continue;
            const double f[3] = {fm[0 * n * n + (long)j * n + i], fm[1 * n * n + (long)j * n + i],
                                 fm[2 * n * n + (long)j * n + i]};
            const double r[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
            const double fv = f[0] * v[0] + f[1] * v[1] + f[2] * v[2];
            for (int a = 0; a < 3; ++a) jpot[a] += r[a] * fv;
        }
        if (stress_ew1p) ew1p_term(stress_ew1p + 9 * i, v, jpot);
    }
    for (int a = 0; a < 3; ++a) jtot[a] = jkin[a] + jpot[a];
}

void ora_heat_flux_tensor(int nreal, const double *w,
                          const double *vx, const double *vy, const double *vz,
                          const double *e_kin, const double *e_pot,
                          const double *stress_ew1p,
                          double jkin[3], double jpot[3], double jtot[3])
{
    memset(jkin, 0, 3 * sizeof(double));
    memset(jpot, 0, 3 * sizeof(double));
    for (int i = 0; i < nreal; ++i) {
        const double v[3] = {vx[i], vy[i], vz[i]};
        const double eng = e_kin[i] + e_pot[i];
        for (int a = 0; a < 3; ++a) jkin[a] += eng * v[a];
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) jpot[b] += v[a] * w[9 * i + 3 * a + b];
        if (stress_ew1p) ew1p_term(stress_ew1p + 9 * i, v, jpot);
    }
    for (int a = 0; a < 3; ++a) jtot[a] = jkin[a] + jpot[a];
}
