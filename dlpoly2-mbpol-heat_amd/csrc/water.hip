// water.hip -- intramolecular and dispersion terms of MB-pol water on gfx950: monomer
// bonds (bndfrc.f), Partridge-Schwenke coupling terms (angfrc.f keya = 9 ->
// ps_type_couple.f90), Tang-Toennies-damped dispersion from interpolation tables
// (forgen.f, srfrce2.f) and the long-range dispersion correction (lrcorrect.f).
//
// GPU organisation (no floating-point atomics; every per-atom sum has one owner):
//   * k_wt_monomer: one wavefront per water.  Lanes 0..47 build the power tables
//     x^k (k = 0..15) of the three PS variables in LDS by the same multiplication chain the
//     reference uses; each lane then takes every 64th of the 244 expansion terms (lanes
//     0..8 also one mass-correction term), and the wave sums energy, H1 and H2 forces and
//     virial with xor shuffles.  Lane 0 adds the three bonds and writes the per-atom force,
//     energy and force-matrix contraction W of O, H1 and H2, and the water's scalars;
//   * k_wt_disp: one wavefront per real atom i owns its full dispersion row (both
//     directions of every pair): lanes stride over all atoms j of other waters, take the
//     minimum image, and interpolate energy and force from the tables (3-point formula of
//     srfrce2.f).  The row sums, plus atom i's monomer terms, give its force, energy and W
//     directly; pair scalars are halved so that each pair counts once;
//   * k_wt_reduce: one block sums the per-water and per-atom scalars in a fixed order.
// The long-range correction depends only on the cell volume, the atom counts and the
// parameters and is evaluated on the host.
#include <math.h>
#include <string.h>

#include <string>

#include "mbgpu_internal.h"
#include "gen/ps_table.inc"

namespace {

constexpr double PS_REOH = 0.958649;               // module.f90:480
constexpr double PS_B1 = 2.0;                      // module.f90:482
constexpr double PS_HHMASS = 1836.152697;          // module.f90:470-471
constexpr double PS_HDMASS = 3670.483031;
constexpr double PS_ENGCNV = 1.19627;              // module.f90:486
constexpr double PS_F5Z = 0.99967788500000, PS_FBASIS = 0.15860145369897;
constexpr double PS_FCORE = -1.6351695982132, PS_FREST = 1.0;
constexpr double PS_EOFFSET = 2.0384848702093006e-6 * 2625.50 / 100.0;   // ps_type_couple.f90:251
constexpr double PI_D = 3.141592653589793;

__device__ __forceinline__ double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// f_n(x) = 1 - exp(-x) sum_{k<=n} x^k / k!, with the reference's series for |f| < 1e-8
// (tang-toennies.f90:5-31)
__host__ __device__ double tt_damp(int n, double x)
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

// bond energy and dV/dr (bndfrc.f:67-108)
__device__ __forceinline__ void bond_fn(int key, const double* p, double r, double& v, double& g)
{
    if (key == 1) {
        v = 0.5 * p[0] * (r - p[1]) * (r - p[1]);
        g = p[0] * (r - p[1]);
    } else if (key == 2) {
        const double e = exp(-p[2] * (r - p[1]));
        v = p[0] * ((1.0 - e) * (1.0 - e) - 1.0);
        g = 2.0 * p[0] * p[2] * (1.0 - e) * e;
    } else if (key == 7) {
        v = p[0] * exp(-p[1] * r);
        g = -p[0] * p[1] * exp(-p[1] * r);
    } else {
        v = 0.0;
        g = 0.0;
    }
}

// vv12 / gg12 (forgen.f:338-357) on the table grid: row t holds r = i * dlrpot, i = 1..mxgrid
__global__ __launch_bounds__(256) void k_wt_table(int mxgrid, int ngrid, double dlrpot, double4 c6a6c8a8_0,
                                                  double4 p1, double4 p2, double* __restrict__ vvv,
                                                  double* __restrict__ ggg)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 3 * ngrid) return;
    const int t = k / ngrid, i = k - t * ngrid;
    const double4 q = t == 0 ? c6a6c8a8_0 : (t == 1 ? p1 : p2);
    double v = 0.0, g = 0.0;
    if (i >= 1 && i <= mxgrid) {
        const double r = (double)i * dlrpot;
        const double c6 = q.x, a6 = q.y, c8 = q.z, a8 = q.w;
        const double f6 = tt_damp(6, a6 * r), f8 = tt_damp(8, a8 * r);
        v = -c6 * f6 / pow(r, 6) - c8 * f8 / pow(r, 8);
        g = -6 * c6 * f6 / pow(r, 6) - 8 * c8 * f8 / pow(r, 8) + r * c6 * pow(a6, 7) * exp(-a6 * r) / 720 +
            r * c8 * pow(a8, 9) * exp(-a8 * r) / 40320;
    }
    vvv[k] = v;
    ggg[k] = g;
}

// One wavefront per water: bonds O-H1, O-H2, H1-H2 and the PS coupling terms.
// mono[i][13] = force (3), energy, W (9) of atom i; wsc[w][10] = engbnd, virbnd, engang,
// virang, stress xx xy xz yy yz zz.
__global__ __launch_bounds__(256) void k_wt_monomer(WaterDev P, const double* __restrict__ x,
                                                    const double* __restrict__ y, const double* __restrict__ z,
                                                    const double* __restrict__ psc, const double* __restrict__ pcm,
                                                    const int* __restrict__ pidx, const int* __restrict__ pidxm,
                                                    double* __restrict__ mono, double* __restrict__ wsc)
{
#pragma clang fp contract(off)
    __shared__ double fm[4][16][3];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int w0 = blockIdx.x * 4 + wv;
    const bool act = w0 < P.nw;                  // whole wavefront; inactive waves still meet the barrier
    const int w = act ? w0 : P.nw - 1;
    const int io = 3 * w, ih1 = io + 1, ih2 = io + 2;
    const double xo = x[io], yo = y[io], zo = z[io];
    const double x1 = x[ih1], y1 = y[ih1], z1 = z[ih1];
    const double x2 = x[ih2], y2 = y[ih2], z2 = z[ih2];
    // PS geometry (angfrc.f:176-240): a = H1, b = O, c = H2
    double dab[3] = {x1 - xo, y1 - yo, z1 - zo}, dbc[3] = {x2 - xo, y2 - yo, z2 - zo};
    image_vec(P.g, dab[0], dab[1], dab[2]);
    image_vec(P.g, dbc[0], dbc[1], dbc[2]);
    const double rab = sqrt(norm2_nc(dab[0], dab[1], dab[2])), rbc = sqrt(norm2_nc(dbc[0], dbc[1], dbc[2]));
    const double rrab = 1.0 / rab, rrbc = 1.0 / rbc;
    const double u1[3] = {dab[0] * rrab, dab[1] * rrab, dab[2] * rrab};
    const double u2[3] = {dbc[0] * rrbc, dbc[1] * rrbc, dbc[2] * rrbc};
    const double theta = acos((u1[0] * u2[0] + u1[1] * u2[1]) + u1[2] * u2[2]);
    const double cost = cos(theta);
    const double cose = cos(104.3475 * PI_D / 180.0);
    const double xm = 1.0 / PS_HHMASS;
    double pt = 0.0, v1 = 0.0, v2 = 0.0, vt = 0.0, fa[3] = {0, 0, 0}, fc[3] = {0, 0, 0};
    // power tables fm[k][v] = x_v^(k-1), fm[0] = 0 (ps_type_couple.f90:64-75): lane k of
    // each 16-lane group repeats the reference's multiplication chain up to its power
    if (lane < 48) {
        const int v = lane >> 4, k = lane & 15;
        const double xv = v == 0 ? (rab - PS_REOH) / PS_REOH : (v == 1 ? (rbc - PS_REOH) / PS_REOH : cost - cose);
        double p = k == 0 ? 0.0 : 1.0;
        for (int q = 2; q <= k; ++q) p = p * xv;
        fm[wv][k][v] = p;
    }
    __syncthreads();
    if (!act) return;
    if (P.ps_monomer) {
        const double(*F)[3] = fm[wv];
        for (int j = 1 + lane; j < PS_NTERMS; j += 64) {
            const int pk = pidx[j];
            const int i1 = pk & 255, i2 = (pk >> 8) & 255, i3 = pk >> 16;
            const double c = psc[j];
            const double dum = F[i1][0] * F[i2][1] + F[i2][0] * F[i1][1];
            pt = pt + c * dum * F[i3][2];
            for (int q = 0; q < 3; ++q) {
                const double d1 = (i1 - 1) * F[i1 - 1][0] * u1[q] / PS_REOH;
                const double d2 = (i2 - 1) * F[i2 - 1][0] * u1[q] / PS_REOH;
                const double d3 = (i3 - 1) * F[i3 - 1][2] * (u2[q] - cost * u1[q]) / rab;
                const double dd = d1 * F[i2][1] * F[i3][2] + d2 * F[i1][1] * F[i3][2] + dum * d3;
                fa[q] = fa[q] - c * dd;
            }
            for (int q = 0; q < 3; ++q) {
                const double d1 = (i2 - 1) * F[i2 - 1][1] * u2[q] / PS_REOH;
                const double d2 = (i1 - 1) * F[i1 - 1][1] * u2[q] / PS_REOH;
                const double d3 = (i3 - 1) * F[i3 - 1][2] * (u1[q] - cost * u2[q]) / rbc;
                const double dd = F[i1][0] * d1 * F[i3][2] + F[i2][0] * d2 * F[i3][2] + dum * d3;
                fc[q] = fc[q] - c * dd;
            }
            vt = vt + c * (rab * F[i2][1] * (i1 - 1) * F[i1 - 1][0] / PS_REOH +
                           rbc * F[i1][0] * (i2 - 1) * F[i2 - 1][1] / PS_REOH +
                           rab * F[i1][1] * (i2 - 1) * F[i2 - 1][0] / PS_REOH +
                           rbc * F[i2][0] * (i1 - 1) * F[i1 - 1][1] / PS_REOH) *
                          F[i3][2];
        }
        if (lane < PS_NMASS) {   // mass-correction terms (ps_type_couple.f90:163-226)
            const int pk = pidxm[lane];
            const int i1 = pk & 255, i2 = (pk >> 8) & 255, i3 = pk >> 16;
            const double c = pcm[lane];
            v1 = c * F[i1][0] * F[i2][1] * F[i3][2];
            v2 = c * F[i2][0] * F[i1][1] * F[i3][2];
            for (int q = 0; q < 3; ++q) {
                const double d1 = (i1 - 1) * F[i1 - 1][0] * u1[q] / PS_REOH;
                const double d2 = (i2 - 1) * F[i2 - 1][0] * u1[q] / PS_REOH;
                const double d3 = (i3 - 1) * F[i3 - 1][2] * (u2[q] - cost * u1[q]) / rab;
                const double corr = xm * (d1 * F[i2][1] * F[i3][2] + F[i1][0] * F[i2][1] * d3) +
                                    xm * (d2 * F[i1][1] * F[i3][2] + F[i2][0] * F[i1][1] * d3);
                fa[q] = fa[q] - c * corr;
            }
            for (int q = 0; q < 3; ++q) {
                const double d1 = (i2 - 1) * F[i2 - 1][1] * u2[q] / PS_REOH;
                const double d2 = (i1 - 1) * F[i1 - 1][1] * u2[q] / PS_REOH;
                const double d3 = (i3 - 1) * F[i3 - 1][2] * (u1[q] - cost * u2[q]) / rbc;
                const double corr = xm * (F[i1][0] * d1 * F[i3][2] + F[i1][0] * F[i2][1] * d3) +
                                    xm * (F[i2][0] * d2 * F[i3][2] + F[i2][0] * F[i1][1] * d3);
                fc[q] = fc[q] - c * corr;
            }
            vt = vt + c * (xm * rab * F[i2][1] * (i1 - 1) * F[i1 - 1][0] / PS_REOH +
                           xm * rbc * F[i1][0] * (i2 - 1) * F[i2 - 1][1] / PS_REOH +
                           xm * rab * F[i1][1] * (i2 - 1) * F[i2 - 1][0] / PS_REOH +
                           xm * rbc * F[i2][0] * (i1 - 1) * F[i1 - 1][1] / PS_REOH) *
                          F[i3][2];
        }
        pt = wsum(pt);
        v1 = wsum(v1);
        v2 = wsum(v2);
        vt = wsum(vt);
        for (int q = 0; q < 3; ++q) {
            fa[q] = wsum(fa[q]);
            fc[q] = wsum(fc[q]);
        }
    }
    if (lane != 0) return;
    double pterm = 0.0, vterm = 0.0;
    if (P.ps_monomer) {   // prefactor (ps_type_couple.f90:228-258)
        pt = pt + xm * v1 + xm * v2;
        const double efac = exp(-PS_B1 * ((rab - PS_REOH) * (rab - PS_REOH) + (rbc - PS_REOH) * (rbc - PS_REOH)));
        pterm = pt * efac + psc[0] + PS_EOFFSET;
        for (int q = 0; q < 3; ++q) {
            fa[q] = fa[q] * efac + 2.0 * PS_B1 * (rab - PS_REOH) * u1[q] * efac * pt;
            fc[q] = fc[q] * efac + 2.0 * PS_B1 * (rbc - PS_REOH) * u2[q] * efac * pt;
        }
        const double dummy1 = -2.0 * PS_B1 * (rab - PS_REOH) * rab * efac;
        const double dummy2 = -2.0 * PS_B1 * (rbc - PS_REOH) * rbc * efac;
        vterm = dummy1 * pt + dummy2 * pt + efac * vt;
    }
    // per-atom accumulators: 0 = O, 1 = H1, 2 = H2
    double f[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, e[3] = {0, 0, 0}, W[3][9];
    for (int a = 0; a < 3; ++a)
        for (int q = 0; q < 9; ++q) W[a][q] = 0.0;
    const double px[3] = {xo, x1, x2}, py[3] = {yo, y1, y2}, pz[3] = {zo, z1, z2};
    double st[6] = {0, 0, 0, 0, 0, 0}, eb = 0.0, vb = 0.0;
    // force-matrix entry: force F on atom a due to atom b (un-imaged positions in W)
    auto entry = [&](int a, int b, const double* Fv) {
        const double r[3] = {px[a] - px[b], py[a] - py[b], pz[a] - pz[b]};
        for (int q = 0; q < 3; ++q) f[a][q] += Fv[q];
        for (int p = 0; p < 3; ++p)
            for (int s = 0; s < 3; ++s) W[a][3 * p + s] += Fv[p] * r[s];
    };
    // bonds (bndfrc.f:136-330): (O,H1), (O,H2) with key 0 params, (H1,H2) with key 1 params
    for (int b = 0; b < 3; ++b) {
        const int t = b == 2 ? 1 : 0;
        if (P.bond_key[t] == 0) continue;
        const int ia = b == 2 ? 1 : 0, ib = b == 2 ? 2 : 1 + b;
        double d[3] = {px[ia] - px[ib], py[ia] - py[ib], pz[ia] - pz[ib]};
        image_vec(P.g, d[0], d[1], d[2]);
        const double r = sqrt(norm2_nc(d[0], d[1], d[2]));
        const double rr = r > 1e-6 ? 1.0 / r : 0.0;
        double omega, gamma;
        bond_fn(P.bond_key[t], P.bond_prm[t], r, omega, gamma);
        gamma = gamma * rr;
        eb += omega;
        vb += gamma * r * r;
        e[ia] += 0.5 * omega;
        e[ib] += 0.5 * omega;
        const double fv[3] = {-gamma * d[0], -gamma * d[1], -gamma * d[2]};
        const double mf[3] = {gamma * d[0], gamma * d[1], gamma * d[2]};
        entry(ia, ib, fv);
        entry(ib, ia, mf);
        st[0] += d[0] * fv[0];
        st[1] += d[0] * fv[1];
        st[2] += d[0] * fv[2];
        st[3] += d[1] * fv[1];
        st[4] += d[1] * fv[2];
        st[5] += d[2] * fv[2];
    }
    if (P.ps_monomer) {   // angfrc.f:370-428, 450-460
        for (int a = 0; a < 3; ++a) e[a] += pterm / 3;
        const double mfa[3] = {-fa[0], -fa[1], -fa[2]}, mfc[3] = {-fc[0], -fc[1], -fc[2]};
        entry(1, 0, fa);
        entry(0, 1, mfa);
        entry(2, 0, fc);
        entry(0, 2, mfc);
        st[0] += rab * u1[0] * fa[0] + rbc * u2[0] * fc[0];
        st[1] += rab * u1[0] * fa[1] + rbc * u2[0] * fc[1];
        st[2] += rab * u1[0] * fa[2] + rbc * u2[0] * fc[2];
        st[3] += rab * u1[1] * fa[1] + rbc * u2[1] * fc[1];
        st[4] += rab * u1[1] * fa[2] + rbc * u2[1] * fc[2];
        st[5] += rab * u1[2] * fa[2] + rbc * u2[2] * fc[2];
    }
    for (int a = 0; a < 3; ++a) {
        double* m = mono + 13L * (io + a);
        for (int q = 0; q < 3; ++q) m[q] = f[a][q];
        m[3] = e[a];
        for (int q = 0; q < 9; ++q) m[4 + q] = W[a][q];
    }
    double* s = wsc + 10L * w;
    s[0] = eb;
    s[1] = vb;
    s[2] = pterm;
    s[3] = vterm;
    for (int q = 0; q < 6; ++q) s[4 + q] = st[q];
}

// One wavefront per real atom: its dispersion row plus its monomer terms.
// asc[i][8] = half pair energy, half pair virial, half pair stress (6).
__global__ __launch_bounds__(256) void k_wt_disp(WaterDev P, const double* __restrict__ x,
                                                 const double* __restrict__ y, const double* __restrict__ z,
                                                 const double* __restrict__ vvv, const double* __restrict__ ggg,
                                                 const double* __restrict__ mono, double* __restrict__ fx,
                                                 double* __restrict__ fy, double* __restrict__ fz,
                                                 double* __restrict__ fw, double* __restrict__ eout,
                                                 double* __restrict__ wout, double* __restrict__ asc)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= P.nr) return;
    const double xi = x[i], yi = y[i], zi = z[i];
    const int wi = i / 3, ti = (i % 3) != 0;
    double a[20];
    for (int q = 0; q < 20; ++q) a[q] = 0.0;   // f3, e, W9, st6, vir
    if (P.dispersion) {
        auto pair = [&](int j) {
#pragma clang fp contract(off)
            const int tp = ti + ((j % 3) != 0);
            if (j / 3 == wi || !((P.tmask >> tp) & 1)) return;
            const double rx = xi - x[j], ry = yi - y[j], rz = zi - z[j];
            double dx = rx, dy = ry, dz = rz;
            image_vec(P.g, dx, dy, dz);
            const double rsq = norm2_nc(dx, dy, dz);
            if (!(P.rc2 > rsq)) return;
            const double rrr = sqrt(rsq);
            const int l = (int)(rrr * P.rdr);
            const double ppp = rrr * P.rdr - (double)l;
            const double* vt = vvv + (long)tp * P.ngrid;
            const double* gt = ggg + (long)tp * P.ngrid;
            double t1 = vt[l] + (vt[l + 1] - vt[l]) * ppp;
            double t2 = vt[l + 1] + (vt[l + 2] - vt[l + 1]) * (ppp - 1.0);
            const double e = t1 + (t2 - t1) * ppp * 0.5;
            t1 = gt[l] + (gt[l + 1] - gt[l]) * ppp;
            t2 = gt[l + 1] + (gt[l + 2] - gt[l + 1]) * (ppp - 1.0);
            const double gamma = (t1 + (t2 - t1) * ppp * 0.5) / rsq;
            const double fv[3] = {gamma * dx, gamma * dy, gamma * dz};
            const double r[3] = {rx, ry, rz}, d[3] = {dx, dy, dz};
            for (int q = 0; q < 3; ++q) a[q] += fv[q];
            a[3] += e;
            for (int p = 0; p < 3; ++p)
                for (int s = 0; s < 3; ++s) a[4 + 3 * p + s] += fv[p] * r[s];
            a[13] += d[0] * fv[0];
            a[14] += d[0] * fv[1];
            a[15] += d[0] * fv[2];
            a[16] += d[1] * fv[1];
            a[17] += d[1] * fv[2];
            a[18] += d[2] * fv[2];
            a[19] -= gamma * rsq;
        };
        if (P.cells.on) {
            // the 27 link cells around atom i (fixed cell order, ascending index within one)
            for_neighbour_cells(P.cells, i, [&](int c0, int c1) {
                for (int k = c0 + lane; k < c1; k += 64) pair(P.cells.idx[k]);
            });
        } else {
            for (int j = lane; j < P.nr; j += 64) pair(j);
        }
        for (int q = 0; q < 20; ++q) a[q] = wsum(a[q]);
    }
    if (lane != 0) return;
    const double* m = mono + 13L * i;
    const double f0 = m[0] + a[0], f1 = m[1] + a[1], f2 = m[2] + a[2];
    fw[i] = f0;
    fw[P.nr + i] = f1;
    fw[2L * P.nr + i] = f2;
    if (fx) fx[i] += f0;
    if (fy) fy[i] += f1;
    if (fz) fz[i] += f2;
    eout[i] = m[3] + 0.5 * a[3];
    for (int q = 0; q < 9; ++q) wout[9L * i + q] = m[4 + q] + a[4 + q];
    double* s = asc + 8L * i;
    s[0] = 0.5 * a[3];
    s[1] = 0.5 * a[19];
    for (int q = 0; q < 6; ++q) s[2 + q] = 0.5 * a[13 + q];
}

// Fixed-order sums of the per-water [nw][10] and per-atom [nr][8] scalars into out[18].
__global__ __launch_bounds__(1024) void k_wt_reduce(int nw, int nr, const double* __restrict__ wsc,
                                                    const double* __restrict__ asc, double* __restrict__ out)
{
    __shared__ double red[16][18];
    double v[18];
    for (int q = 0; q < 18; ++q) v[q] = 0.0;
    for (int w = threadIdx.x; w < nw; w += blockDim.x)
        for (int q = 0; q < 10; ++q) v[q] += wsc[10L * w + q];
    for (int i = threadIdx.x; i < nr; i += blockDim.x)
        for (int q = 0; q < 8; ++q) v[10 + q] += asc[8L * i + q];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = 0; q < 18; ++q) {
        const double s = wsum(v[q]);
        if (lane == 0) red[wv][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < 18) {
        double s = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

template <typename T>
hipError_t dalloc(T** p, size_t n)
{
    return mbgpu_alloc(p, n);
}

template <typename T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// init_ps_type_pes (module.f90:513-600): scale factors, mass correction, unit conversion
void ps_process(double* c5z, double* cmass)
{
    double crest[PS_NTERMS];
    for (int j = 0; j < PS_NTERMS; ++j) crest[j] = ps_cols[j][3];
    for (int i = 0; i < PS_NMASS; ++i) {
        cmass[i] = ps_cmass0[i] / (1.0 / PS_HDMASS - 1.0 / PS_HHMASS);
        double fac_corr = cmass[i] / PS_HHMASS;
        if (ps_idxm[i][0] == ps_idxm[i][1]) fac_corr = fac_corr * 0.5;
        for (int j = 0; j < PS_NTERMS; ++j)
            if (ps_idx[j][0] == ps_idxm[i][0] && ps_idx[j][1] == ps_idxm[i][1] && ps_idx[j][2] == ps_idxm[i][2]) {
                crest[j] = crest[j] - fac_corr;
                break;
            }
        for (int j = 0; j < PS_NTERMS; ++j)
            if (ps_idx[j][1] == ps_idxm[i][0] && ps_idx[j][0] == ps_idxm[i][1] && ps_idx[j][2] == ps_idxm[i][2]) {
                crest[j] = crest[j] - fac_corr;
                break;
            }
    }
    for (int j = 0; j < PS_NTERMS; ++j)
        c5z[j] = (PS_F5Z * ps_cols[j][0] + PS_FBASIS * ps_cols[j][1] + PS_FCORE * ps_cols[j][2] + PS_FREST * crest[j]) *
                 PS_ENGCNV;
    for (int i = 0; i < PS_NMASS; ++i) cmass[i] = cmass[i] * PS_FREST * PS_ENGCNV;
    c5z[0] = c5z[0] * 2.0;
}

// tang_toennies_lr (tang-toennies.f90:33-74), whose "factorial" multiplies n by itself n times
double tt_lr(int n, double x)
{
    double fact = 1.0;
    for (int i = n; i >= 1; --i) fact = fact * n;
    const double r = tt_damp(n, x) / pow(x, n - 3) + exp(-x) * (x * x * (3.0 + x) + 6 * (1.0 + x)) / fact;
    return r / (n - 3);
}

// elrc and virlrc (lrcorrect.f:100-209, el12 / pl12 :270-302) for nw O and 2nw H atoms
void long_range(const mbgpu_water_params* p, int nw, double volm, double& elrc, double& virlrc)
{
    const double num[2] = {(double)nw, 2.0 * nw};
    double e = 0.0, pl = 0.0;
    for (int ti = 0; ti < 2; ++ti)
        for (int tj = 0; tj <= ti; ++tj) {
            const int t = ti + tj;
            const double r = p->rvdw, c6 = p->c6[t], a6 = p->a6[t], c8 = p->c8[t], a8 = p->a8[t];
            const double el = -c6 * pow(a6, 3) * tt_lr(6, a6 * r) - c8 * pow(a8, 5) * tt_lr(8, a8 * r);
            const double pp = 3 * el - c6 * tt_damp(6, a6 * r) / pow(r, 3) - c8 * tt_damp(8, a8 * r) / pow(r, 5);
            double eadd = el, padd = pp;
            if (ti != tj) {
                eadd = eadd * 2.0;
                padd = padd * 2.0;
            }
            const double denprd = 2.0 * PI_D * (num[ti] * num[tj]) / (volm * volm);
            e = e + volm * denprd * eadd;
            pl = pl + denprd * padd / 3.0;
        }
    elrc = e;
    virlrc = pl * (-3.0 * volm);
}

}  // namespace

void water_release(WaterState& S)
{
    WaterBuffers& B = S.B;
    dfree(B.mono); dfree(B.wsc); dfree(B.asc); dfree(B.fw); dfree(B.eout); dfree(B.wout);
    dfree(B.vvv); dfree(B.ggg); dfree(B.psc); dfree(B.pcm); dfree(B.pidx); dfree(B.pidxm); dfree(B.scal);
    cells_free(S.cells);
    S.nw = 0;
    S.have_tab = false;
    S.ready = false;
}

#define WCHK(x)                                                                        \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            err = std::string(#x) + ": " + hipGetErrorString(e_);                      \
            return (int)e_;                                                            \
        }                                                                              \
    } while (0)

int water_configure(WaterState& S, int nw, const CellGeom& g, const double* cell, const mbgpu_water_params* p,
                    hipStream_t s, std::string& err)
{
    for (int t = 0; t < 2; ++t) {
        const int k = p->bond_key[t];
        if (k != 0 && k != 1 && k != 2 && k != 7) {
            err = "bond_key[" + std::to_string(t) + "] = " + std::to_string(k) +
                  " is not one of 0 (none), 1 (harm), 2 (mors), 7 (expo)";
            return -1;
        }
    }
    if (p->dispersion || p->long_range) {
        if (!(p->rvdw > 0.0) || !(p->rcut >= p->rvdw)) {
            err = "need 0 < rvdw <= rcut";
            return -1;
        }
        for (int t = 0; t < 3; ++t)
            if ((p->c6[t] != 0.0 && !(p->a6[t] > 0.0)) || (p->c8[t] != 0.0 && !(p->a8[t] > 0.0)) ||
                (p->long_range && (!(p->a6[t] > 0.0) || !(p->a8[t] > 0.0)))) {
                err = "dispersion damping parameters a6, a8 must be positive (the long-range correction "
                      "divides by powers of a*rvdw)";
                return -1;
            }
    }
    WaterBuffers& B = S.B;
    const int nr = 3 * nw;
    if (S.nw != nw) {
        water_release(S);
        WCHK(dalloc(&B.mono, (size_t)13 * nr));
        WCHK(dalloc(&B.wsc, (size_t)10 * nw));
        WCHK(dalloc(&B.asc, (size_t)8 * nr));
        WCHK(dalloc(&B.fw, (size_t)3 * nr));
        WCHK(dalloc(&B.eout, (size_t)nr));
        WCHK(dalloc(&B.wout, (size_t)9 * nr));
        WCHK(dalloc(&B.psc, PS_NTERMS));
        WCHK(dalloc(&B.pcm, PS_NMASS));
        WCHK(dalloc(&B.pidx, PS_NTERMS));
        WCHK(dalloc(&B.pidxm, PS_NMASS));
        WCHK(dalloc(&B.scal, 24));
        double c5z[PS_NTERMS], cm[PS_NMASS];
        int pk[PS_NTERMS], pkm[PS_NMASS];
        ps_process(c5z, cm);
        for (int j = 0; j < PS_NTERMS; ++j) pk[j] = ps_idx[j][0] | (ps_idx[j][1] << 8) | (ps_idx[j][2] << 16);
        for (int j = 0; j < PS_NMASS; ++j) pkm[j] = ps_idxm[j][0] | (ps_idxm[j][1] << 8) | (ps_idxm[j][2] << 16);
        WCHK(hipMemcpy(B.psc, c5z, sizeof c5z, hipMemcpyHostToDevice));
        WCHK(hipMemcpy(B.pcm, cm, sizeof cm, hipMemcpyHostToDevice));
        WCHK(hipMemcpy(B.pidx, pk, sizeof pk, hipMemcpyHostToDevice));
        WCHK(hipMemcpy(B.pidxm, pkm, sizeof pkm, hipMemcpyHostToDevice));
        S.nw = nw;
    }
    WaterDev& P = S.P;
    P.nw = nw;
    P.nr = nr;
    P.g = g;
    for (int t = 0; t < 2; ++t) {
        P.bond_key[t] = p->bond_key[t];
        for (int q = 0; q < 3; ++q) P.bond_prm[t][q] = p->bond_prm[t][q];
    }
    P.ps_monomer = p->ps_monomer != 0;
    P.dispersion = p->dispersion != 0;
    P.tmask = 0;
    if (P.dispersion) {
        int mxgrid = (int)(p->rcut / 0.0001 + 0.5) + 4;
        if (mxgrid < 1000) mxgrid = 1000;
        const double key[14] = {p->rcut, p->rvdw, p->c6[0], p->c6[1], p->c6[2], p->a6[0], p->a6[1],
                                p->a6[2], p->c8[0], p->c8[1], p->c8[2], p->a8[0], p->a8[1], p->a8[2]};
        const double dlrpot = p->rvdw / (double)(mxgrid - 4);
        if (!S.have_tab || memcmp(key, S.tab_key, sizeof key) != 0) {
            if (B.ngrid != mxgrid + 3) {
                dfree(B.vvv);
                dfree(B.ggg);
                B.ngrid = mxgrid + 3;
                WCHK(dalloc(&B.vvv, (size_t)3 * B.ngrid));
                WCHK(dalloc(&B.ggg, (size_t)3 * B.ngrid));
            }
            const int nk = 3 * B.ngrid;
            hipLaunchKernelGGL(k_wt_table, dim3((nk + 255) / 256), dim3(256), 0, s, mxgrid, B.ngrid, dlrpot,
                               make_double4(p->c6[0], p->a6[0], p->c8[0], p->a8[0]),
                               make_double4(p->c6[1], p->a6[1], p->c8[1], p->a8[1]),
                               make_double4(p->c6[2], p->a6[2], p->c8[2], p->a8[2]), B.vvv, B.ggg);
            WCHK(hipGetLastError());
            memcpy(S.tab_key, key, sizeof key);
            S.have_tab = true;
        }
        P.ngrid = B.ngrid;
        P.rc2 = p->rvdw * p->rvdw;
        P.rdr = 1.0 / dlrpot;
        for (int t = 0; t < 3; ++t)
            if (p->c6[t] != 0.0 || p->c8[t] != 0.0) P.tmask |= 1 << t;
    }
    // link cells of the real atoms for the dispersion rows (the test is rsq < rvdw^2)
    P.cells.on = 0;
    if (p->dispersion) {
        S.cells.V.on = 0;
        if (g.imcon != 0 && cells_plan(S.cells, cell, p->rvdw, 3 * nw, CELLS_MIN_ATOMS, err)) return -1;
        P.cells = S.cells.V;
    }
    // long-range correction on the host
    S.out[6] = S.out[7] = 0.0;
    if (p->long_range && g.imcon != 0) {   // lrcorrect.f:114: none without a periodic cell
        const double* a = cell;
        const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[3] * (a[1] * a[8] - a[2] * a[7]) +
                           a[6] * (a[1] * a[5] - a[2] * a[4]);
        long_range(p, nw, fabs(det), S.out[6], S.out[7]);
    }
    return 0;
}

// The device work of an evaluation (kernels only): monomers, link cells, dispersion rows,
// reductions into B.scal.
int water_enqueue(WaterState& S, const double* x, const double* y, const double* z, double* fx, double* fy,
                  double* fz, hipStream_t s)
{
    WaterBuffers& B = S.B;
    const WaterDev& P = S.P;
    hipLaunchKernelGGL(k_wt_monomer, dim3((P.nw + 3) / 4), dim3(256), 0, s, P, x, y, z, B.psc, B.pcm, B.pidx,
                       B.pidxm, B.mono, B.wsc);
    if (P.cells.on) {
        const int rc = cells_build(S.cells, P.nr, x, y, z, 1, s);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_wt_disp, dim3((P.nr + 3) / 4), dim3(256), 0, s, P, x, y, z, B.vvv, B.ggg, B.mono, fx, fy,
                       fz, B.fw, B.eout, B.wout, B.asc);
    hipLaunchKernelGGL(k_wt_reduce, dim3(1), dim3(1024), 0, s, P.nw, P.nr, B.wsc, B.asc, B.scal);
    return (int)hipGetLastError();
}

int water_evaluate(WaterState& S, const double* x, const double* y, const double* z, double* fx, double* fy,
                   double* fz, hipStream_t s, std::string& err)
{
    WaterBuffers& B = S.B;
    S.ready = false;
    WCHK((hipError_t)water_enqueue(S, x, y, z, fx, fy, fz, s));
    double r[18];
    WCHK(hipMemcpyAsync(r, B.scal, sizeof r, hipMemcpyDeviceToHost, s));
    WCHK(hipStreamSynchronize(s));
    water_finish(S, r);
    return 0;
}

// The host side of an evaluation's end: S.out and S.stress from the 18 reduced sums.
void water_finish(WaterState& S, const double r[18])
{
    S.out[0] = r[0];
    S.out[1] = r[1];
    S.out[2] = r[2];
    S.out[3] = r[3];
    S.out[4] = r[10];
    S.out[5] = r[11];
    const double st[6] = {r[4] + r[12], r[5] + r[13], r[6] + r[14], r[7] + r[15], r[8] + r[16], r[9] + r[17]};
    const double full[9] = {st[0], st[1], st[2], st[1], st[3], st[4], st[2], st[4], st[5]};
    for (int q = 0; q < 9; ++q) S.stress[q] = full[q];
    S.ready = true;
}
