/*
 * ring_oracle.c -- CPU restatement of the reference's in-process ring-polymer pieces:
 * normal-mode matrices, the bead transforms and the spring forces.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restated (reference file:line):
 *   normal-mode matrices and eigenvalues   centroid_matrix.f:3-135 (transformation_matrix)
 *   cartesian <-> normal-mode transforms    centroid_matrix.f:140-476
 *   spring forces, energy, centroid-virial  centroid_spring_force.f:3-157
 *   constants hbar, kB                      centroid_module.f90:19-21
 *
 * Layout: bead-major, a bead's coordinates as [3][natom] (x block, y block, z block), so
 * one bead is one DL_POLY configuration (xxx, yyy, zzz).  Matrices are P x P row-major with
 * element [k][b] = M(k+1, b+1) of the reference's Fortran arrays.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define PI_ 3.141592653589793

/* transformation_matrix (centroid_matrix.f:3-135).  langevin = 0: the Cao-Voth ordering
 * with nmode_to_cart(:,1) = 1 and cart_to_nmode = C^T / P; langevin = 1: the orthonormal
 * ordering of the PIGLET branch.  lambda: normal-mode eigenvalues (mass factors). */
void ora_ring_matrices(int P, int langevin, double *n2c, double *c2n, double *tc2n, double *lambda)
{
    memset(n2c, 0, sizeof(double) * P * P);
#define C_(n1, n2) n2c[((n1) - 1) * P + ((n2) - 1)]
    for (int n1 = 1; n1 <= P; ++n1) C_(n1, 1) = 1.0;
    for (int n1 = 1; n1 <= P / 2; ++n1) {
        C_(2 * n1, P) = -1.0;
        C_(2 * n1 - 1, P) = 1.0;
    }
    for (int n1 = 1; n1 <= P; ++n1) {
        const double t = 2.0 * PI_ * (double)(n1 - 1) / (double)P;
        for (int n2 = 1; n2 <= P / 2 - 1; ++n2) {
            C_(n1, 2 * n2) = sqrt(2.0) * cos(t * (double)n2);
            C_(n1, 2 * n2 + 1) = -sqrt(2.0) * sin(t * (double)n2);
        }
    }
    for (int n2 = 1; n2 <= P; ++n2)
        for (int n1 = 1; n1 <= P; ++n1) {
            c2n[(n2 - 1) * P + (n1 - 1)] = C_(n1, n2) / (double)P;
            tc2n[(n2 - 1) * P + (n1 - 1)] = C_(n1, n2);
        }
    if (langevin) {
        for (int n1 = 1; n1 <= P; ++n1) {
            double f = sqrt(1.0 / P);
            C_(n1, 1) = f;
            C_(n1, P / 2 + 1) = f * ((n1 % 2) ? -1.0 : 1.0);
            f = sqrt(2.0 / P);
            const double t = 2.0 * PI_ * (double)n1 / (double)P;
            for (int n2 = 2; n2 <= P / 2; ++n2) {
                C_(n1, n2) = f * cos(t * (double)(n2 - 1));
                C_(n1, n2 + P / 2) = f * sin(t * ((double)n2 + (double)P / 2.0 - 1.0));
            }
        }
        for (int n2 = 1; n2 <= P; ++n2)
            for (int n1 = 1; n1 <= P; ++n1) {
                c2n[(n2 - 1) * P + (n1 - 1)] = C_(n1, n2);
                tc2n[(n2 - 1) * P + (n1 - 1)] = C_(n1, n2);
            }
        for (int n = 0; n < P; ++n) lambda[n] = 1.0;
    } else {
        lambda[0] = 0.0;
        for (int n = 1; n <= P / 2 - 1; ++n) {
            double t = 2.0 * PI_ * (double)n / (double)P;
            t = 2.0 * (1.0 - cos(t));
            lambda[2 * n - 1] = t * (double)P;
            lambda[2 * n] = t * (double)P;
        }
        lambda[P - 1] = 4.0 * (double)P;
    }
#undef C_
}

/* out[k] = sum_b M[k][b] in[b] per coordinate (centroid_matrix.f:140-476: every transform
 * is this product with cart_to_nmode, tcart_to_tnmode or nmode_to_cart) */
void ora_ring_transform(int P, int ncoord, const double *M, const double *in, double *out)
{
    for (int k = 0; k < P; ++k)
        for (int c = 0; c < ncoord; ++c) {
            double s = 0.0;
            for (int b = 0; b < P; ++b) s += M[k * P + b] * in[(long)b * ncoord + c];
            out[(long)k * ncoord + c] = s;
        }
}

/* centroid_spring_force.f:41-157 for every bead: coeff = P / (hbar beta)^2; with the
 * centroid xc ([3][natom], the mode-1 coordinates times fac1), d1 = x_k - xc,
 * d2 = x_k - x_{k-1}, d3 = x_k - x_{k+1} (ring order);
 *   Epot_deriv  = 0.5 * sum_{k,i} (-d1 . f_k,i) / force_scale   (f before the springs)
 *   Epot_spring = 0.5 * coeff * sum_{k,i} m_i |d2|^2
 *   f_k,i      -= coeff m_i (d2 + d3) fac3 */
void ora_ring_spring(int P, int natom, double coeff, double fac3, double force_scale, const double *mass,
                     const double *x, const double *xc, double *f, double *epot_spring, double *epot_deriv)
{
    double se = 0.0, sd = 0.0;
    const long nb = 3L * natom;
    for (int k = 0; k < P; ++k) {
        const int kp = (k + P - 1) % P, kn = (k + 1) % P;
        for (int i = 0; i < natom; ++i) {
            double d1[3], d2[3], d3[3];
            for (int a = 0; a < 3; ++a) {
                const double xk = x[k * nb + (long)a * natom + i];
                d1[a] = xk - xc[(long)a * natom + i];
                d2[a] = xk - x[kp * nb + (long)a * natom + i];
                d3[a] = xk - x[kn * nb + (long)a * natom + i];
            }
            double *fk = f + k * nb;
            sd = sd - d1[0] * fk[i] - d1[1] * fk[natom + i] - d1[2] * fk[2L * natom + i];
            se = se + mass[i] * (d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2]);
            for (int a = 0; a < 3; ++a) fk[(long)a * natom + i] -= coeff * mass[i] * (d2[a] + d3[a]) * fac3;
        }
    }
    *epot_spring = 0.5 * coeff * se;
    *epot_deriv = 0.5 * sd / force_scale;
}
