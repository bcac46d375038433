// elec_kernels.hip -- polarizable Ewald electrostatics of MB-pol water on gfx950.
//
// The reference computes, once per force evaluation (forces.f:268-404 with the
// subroutines it calls):  M sites and DMS charges (dlpoly_cmd.f:2041-2057,
// qdetermine.f90), the induced dipoles by SOR over real-space (rfield.f) and
// reciprocal-space (ewald1p.f, iflag = 0) fields (iter.f), then the polarization energy
// (coulomb_polar.f), reciprocal energy / forces / per-atom stress (ewald1p.f, iflag = 1),
// real-space pair terms (ewald2p.f), intramolecular corrections (ewald3p.f) and the
// charge-derivative forces (qdforce.f90).
//
// GPU organisation (one context, one stream, no atomics on floating-point data: every
// per-site sum is owned by one thread or one wavefront and summed in a fixed order):
//   * real space: full neighbour rows per site (both directions of each pair), so that a
//     wavefront owns a site's row; the Thole-damped dipole tensors of the polarizable
//     pairs are built once per evaluation and reused by every solver iteration;
//   * reciprocal space: the k vectors are grouped into runs of consecutive n at fixed
//     (l, m); exp(i k.r_j) comes from per-site tables of exp(i l a*.r_j) (a* a reciprocal
//     vector) and a complex recurrence along n.  Structure factors are k-parallel with a
//     site split (partials summed in a fixed order); fields and forces are site-parallel
//     with a k split (likewise);
//   * the solver loop runs on the device; a done flag set by the convergence check turns
//     the remaining kernels of a batch of iterations into no-ops.
#include <math.h>

#include "mbgpu_internal.h"
#include "gen/dms_table.inc"

#define R4PIE0 138935.4835            // dl_params.inc
#define SQRPI 1.7724538509055159      // dl_params.inc
#define TWOPI (2.0 * 3.141592653589793)
#define EL_RUN 8                      // k vectors per run (consecutive n)
#define EL_SCHUNK 64                  // sites per chunk of the structure-factor pass

namespace {

__device__ __forceinline__ double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// fixed-order block reduction of NV values per thread into out[0..NV) (blockDim <= 1024)
template <int NV>
__device__ void block_sum_store(double (&v)[NV], double* out)
{
    __shared__ double red[16][NV];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = (blockDim.x + 63) >> 6;
    for (int q = 0; q < NV; ++q) {
        const double s = wsum(v[q]);
        if (lane == 0) red[wv][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int w = 0; w < nwv; ++w) s += red[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// Sum over chunks of in[c * cs + off] (c = 0..nc-1) for the 64 elements of a 256-thread
// block, one element per lane: wave w adds the chunks [w nc / 4, (w + 1) nc / 4) in order
// (coalesced loads of 64 consecutive elements), and the four wave sums are added in wave
// order.  Every thread of the block must call it; the sum is returned on threads 0..63.
__device__ double chunk_sum64(const double* __restrict__ in, long cs, int nc, long off, bool valid)
{
    __shared__ double part[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c0 = (nc * wv) / 4, c1 = (nc * (wv + 1)) / 4;
    double sum = 0.0;
    if (valid) {
        const double* p = in + off;
        int c = c0;
        for (; c + 4 <= c1; c += 4) {
            const double u0 = p[c * cs], u1 = p[(c + 1) * cs], u2 = p[(c + 2) * cs], u3 = p[(c + 3) * cs];
            sum += u0;
            sum += u1;
            sum += u2;
            sum += u3;
        }
        for (; c < c1; ++c) sum += p[c * cs];
    }
    part[wv][lane] = sum;
    __syncthreads();
    const double t = wv == 0 ? ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane] : 0.0;
    __syncthreads();
    return t;
}

// Stopping test and scalars of the conjugate-gradient solver at the start of iteration `it`,
// from the block partials of its last vector update: nb blocks of (r.u, w.u,
// |polr2 r / r4pie0|^2, dipole count, max |m_i|).  Every wavefront that calls it sums the partials in
// the same order and takes lane 0's sums, so all of a launch's wavefronts reach the same
// decision without a grid-wide synchronization.  With `record`, lane 0 stores the iteration
// count, the RMS step and then the done flag or the scalars of the update: beta = g / g_prev,
// alpha = g / (d - beta g / alpha_prev) (first iteration: beta = 0, alpha = g / d), with
// g = r.u and d = w.u (the Chronopoulos-Gear recurrence), and a power of two of about
// 1 / max |m_i| for the particle mesh's spreading of m.  state = [iterations, rms step, done, g,
// alpha, beta, -, d, -, m scale].  Returns true when the dipoles have converged.
__device__ bool cg_step(const ElecDev& P, int it, int nb, const double* __restrict__ blk, double* __restrict__ state,
                        bool record)
{
    const int lane = threadIdx.x & 63;
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0, e = 0.0;
    for (int k = lane; k < nb; k += 64) {
        a += blk[5 * k];
        b += blk[5 * k + 1];
        c += blk[5 * k + 2];
        d += blk[5 * k + 3];
        e = fmax(e, blk[5 * k + 4]);
    }
    a = __shfl(wsum(a), 0, 64);
    b = __shfl(wsum(b), 0, 64);
    c = __shfl(wsum(c), 0, 64);
    d = __shfl(wsum(d), 0, 64);
    for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_xor(e, o, 64));
    const double rms = (1.0 - P.omega) * sqrt(c / (d == 0.0 ? 1.0 : d));
    const bool done = rms <= P.toler;
    if (record && lane == 0) {
        state[0] = it;
        state[1] = rms;
        state[9] = (e > 0.0 && isfinite(e)) ? ldexp(1.0, -(ilogb(e) + 1)) : 1.0;
        if (done) {
            state[2] = 1.0;
        } else {
            const double beta = it ? a / state[3] : 0.0;
            state[4] = it ? a / (b - beta * a / state[4]) : a / b;
            state[5] = beta;
            state[3] = a;
            state[7] = b;
        }
    }
    return done;
}

__device__ __forceinline__ int water_of(const ElecDev& P, int i) { return i < P.nreal ? i / 3 : i - P.nreal; }

__device__ __forceinline__ double site_polr(const ElecDev& P, int i)
{
    return i >= P.nreal ? P.polr[2] : ((i % 3) == 0 ? P.polr[0] : P.polr[1]);
}
__device__ __forceinline__ double site_polr2(const ElecDev& P, int i)
{
    return i >= P.nreal ? P.polr2[2] : ((i % 3) == 0 ? P.polr2[0] : P.polr2[1]);
}

// erfcc.f: Numerical Recipes erfc with single-precision literals promoted to double.
__device__ double erfcc_nr(double x)
{
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
    return x < 0.0 ? 2.0 - r : r;
}

// gammq.f: gammln, gser, gcf, gammq (EPS, FPMIN single-precision values)
__device__ double gammln_nr(double xx)
{
    const double cof[6] = {76.18009172947146, -86.50532032941677, 24.01409824083091,
                           -1.231739572450155, .1208650973866179e-2, -.5395239384953e-5};
    double x = xx, y = x;
    double tmp = x + 5.5;
    tmp = (x + 0.5) * log(tmp) - tmp;
    double ser = 1.000000000190015;
    for (int j = 0; j < 6; ++j) {
        y = y + 1.0;
        ser = ser + cof[j] / y;
    }
    return tmp + log(2.5066282746310005 * ser / x);
}

__device__ double gammq_nr(double a, double x, double gln)
{
    const double EPS = (double)3.e-7f, FPMIN = (double)1.e-30f;
    if (x < a + 1.0) {
        if (x <= 0.0) return 1.0;
        double ap = a, sum = 1.0 / a, del = sum;
        for (int n = 1; n <= 800; ++n) {
            ap = ap + 1.0;
            del = del * x / ap;
            sum = sum + del;
            if (fabs(del) < fabs(sum) * EPS) break;
        }
        return 1.0 - sum * exp(-x + a * log(x) - gln);
    }
    // the continued fraction is multiplied by exp(lf): below -750 that factor is exactly 0 in
    // double precision (underflow), and so is the result whatever the fraction (finite for
    // x >= a + 1), so the fraction is not evaluated
    const double lf = -x + a * log(x) - gln;
    if (lf < -750.0) return 0.0;
    double b = x + 1.0 - a, c = 1.0 / FPMIN, d = 1.0 / b, h = d;
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
    return exp(lf) * h;
}

// 3-point interpolation of the B-function table (rfield.f:102-120, ewald2p.f:160-178)
__device__ __forceinline__ void table_interp(const double* __restrict__ t, double drewd, double rr, double* bbb)
{
    const int ll = (int)(rr / drewd);
    const double ppp = rr / drewd - (double)ll;
    const double* r0 = t + 4 * ll;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const double vk = r0[n], vk1 = r0[4 + n], vk2 = r0[8 + n];
        const double t1 = vk + (vk1 - vk) * ppp;
        const double t2 = vk1 + (vk2 - vk1) * (ppp - 1.0);
        bbb[n] = t1 + (t2 - t1) * ppp * 0.5;
    }
}

// B functions computed directly (excluded pairs, rfield.f:365-374, ewald3p.f:100-110)
__device__ __forceinline__ void bfun_direct(double alpha, double rr, double rsqi, double* bbb)
{
    if (alpha < 1.0e-6) {   // ewald3p.f: "if ewald parameter is too small then turn it off"
        bbb[0] = bbb[1] = bbb[2] = bbb[3] = 0.0;
        return;
    }
    const double a = alpha * rr;
    bbb[0] = (erfcc_nr(a) - 1.0) / rr;
    const double exp2a = exp(-a * a);
    const double c = 2.0 * alpha * alpha;
    double cm = 1.0;
#pragma unroll
    for (int m = 1; m <= 3; ++m) {
        cm *= c;
        bbb[m] = rsqi * ((2.0 * m - 1.0) * bbb[m - 1] + cm * exp2a / alpha / SQRPI);
    }
}

// imaged separation of sites i, j in the reference's arithmetic
__device__ __forceinline__ double sep(const ElecDev& P, const double* __restrict__ x, const double* __restrict__ y,
                                      const double* __restrict__ z, int i, int j, double* d)
{
#pragma clang fp contract(off)
    double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
    image_vec(P.g, dx, dy, dz);
    d[0] = dx;
    d[1] = dy;
    d[2] = dz;
    return norm2_nc(dx, dy, dz);
}

}  // namespace

// ====================================================================== set-up ======

// M sites (dlpoly_cmd.f:2041-2057) and DMS charges with their derivatives
// (qdetermine.f90:118-259; new charge assignment).  One thread per water.
__global__ void k_el_sites(ElecDev P, double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
                           double* __restrict__ chge, double* __restrict__ dg)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= P.nw) return;
    const int io = 3 * w, m = P.nreal + w;
    double r1v[3], r2v[3];
    {
#pragma clang fp contract(off)
        double d1x = x[io + 1] - x[io], d1y = y[io + 1] - y[io], d1z = z[io + 1] - z[io];
        double d2x = x[io + 2] - x[io], d2y = y[io + 2] - y[io], d2z = z[io + 2] - z[io];
        image_vec(P.g, d1x, d1y, d1z);
        image_vec(P.g, d2x, d2y, d2z);
        x[m] = x[io] + (d1x + d2x) * P.gamma / 2.0;
        y[m] = y[io] + (d1y + d2y) * P.gamma / 2.0;
        z[m] = z[io] + (d1z + d2z) * P.gamma / 2.0;
        r1v[0] = d1x;
        r1v[1] = d1y;
        r1v[2] = d1z;
        r2v[0] = d2x;
        r2v[1] = d2y;
        r2v[2] = d2z;
    }
    const double r1 = sqrt(r1v[0] * r1v[0] + r1v[1] * r1v[1] + r1v[2] * r1v[2]);
    const double r2 = sqrt(r2v[0] * r2v[0] + r2v[1] * r2v[1] + r2v[2] * r2v[2]);
    const double cost = (r1v[0] * r2v[0] + r1v[1] * r2v[1] + r1v[2] * r2v[2]) / (r1 * r2);
    // reference geometry of the expansion: ps_type_pes reoh, thetae (module.f90:480-481)
    const double reoh = 0.958649, cose = cos(104.3475 * 3.141592653589793 / 180.0);
    const double x1 = (r1 - reoh) / reoh, x2 = (r2 - reoh) / reoh, x3 = cost - cose;
    double f1[8], f2[8], f3[8];
    f1[0] = f2[0] = f3[0] = 0.0;
    f1[1] = f2[1] = f3[1] = 1.0;
#pragma unroll
    for (int j = 2; j < 8; ++j) {
        f1[j] = f1[j - 1] * x1;
        f2[j] = f2[j - 1] * x2;
        f3[j] = f3[j - 1] * x3;
    }
    double p1 = 0.0, p2 = 0.0, g11 = 0.0, g12 = 0.0, g21 = 0.0, g22 = 0.0, g1c = 0.0, g2c = 0.0;
#pragma unroll
    for (int j = 1; j < DMS_NTERMS; ++j) {
        const int a = dms_idx[j][0], b = dms_idx[j][1], c = dms_idx[j][2];
        const double cf = dms_coef[j];
        p1 = p1 + cf * f1[a] * f2[b] * f3[c];
        p2 = p2 + cf * f1[b] * f2[a] * f3[c];
        g11 = g11 + cf * (a - 1) * f1[a - 1] * f2[b] * f3[c] / reoh;
        g12 = g12 + cf * (b - 1) * f1[a] * f2[b - 1] * f3[c] / reoh;
        g21 = g21 + cf * (b - 1) * f1[b - 1] * f2[a] * f3[c] / reoh;
        g22 = g22 + cf * (a - 1) * f1[b] * f2[a - 1] * f3[c] / reoh;
        g1c = g1c + cf * (c - 1) * f1[a] * f2[b] * f3[c - 1];
        g2c = g2c + cf * (c - 1) * f1[b] * f2[a] * f3[c - 1];
    }
    // DMS parameters (module.f90:818-826)
    const double b1dms = 1.0, adms = 0.2999 * 0.529177249, bdms = -0.6932;
    const double c0 = 1.0099, c1 = -0.1801, c2 = 0.0892;
    const double pl1 = cost, pl2 = 0.5 * (3.0 * pl1 * pl1 - 1.0);
    const double damp = exp(-b1dms * ((r1 - reoh) * (r1 - reoh) + (r2 - reoh) * (r2 - reoh)));
    const double ddampdr1 = -2.0 * b1dms * (r1 - reoh) * damp;
    const double ddampdr2 = -2.0 * b1dms * (r2 - reoh) * damp;
    const double leg = c0 + pl1 * c1 + pl2 * c2;
    const double pr1 = pow(r1, bdms), pr2 = pow(r2, bdms);
    const double pc0 = adms * (pr1 + pr2) * leg;
    const double dpc0dr1 = adms * (bdms * pow(r1, bdms - 1)) * leg;
    const double dpc0dr2 = adms * (bdms * pow(r2, bdms - 1)) * leg;
    const double dpc0dcos = adms * (pr1 + pr2) * (c1 + 3.0 * c2 * pl1);
    double* G = dg + 6 * w;
    G[0] = g11 * damp + p1 * ddampdr1 + dpc0dr1;
    G[1] = g12 * damp + p1 * ddampdr2 + dpc0dr2;
    G[2] = g21 * damp + p2 * ddampdr1 + dpc0dr1;
    G[3] = g22 * damp + p2 * ddampdr2 + dpc0dr2;
    G[4] = g1c * damp + dpc0dcos;
    G[5] = g2c * damp + dpc0dcos;
    p1 = dms_coef[0] + p1 * damp + pc0;
    p2 = dms_coef[0] + p2 * damp + pc0;
    const double factor = P.gamma / 2.0 / (1.0 - P.gamma);
    const double q1 = p1 + factor * (p1 + p2), q2 = p2 + factor * (p1 + p2);
    chge[io] = P.q_o;
    chge[io + 1] = q1;
    chge[io + 2] = q2;
    chge[m] = -(q1 + q2);
}

// Neighbour rows: every site of another water within rcut (superset by a relative 1e-9 of
// rcut^2; each kernel applies the reference's own test), found in the 27 link cells around
// the site when the box holds three cells per axis (the Verlet list over link cells of
// parlink.f), else by a scan of every site; either way in an order fixed by the positions.
// Row P holds the partners of a polarizable site that are polarizable.  One wavefront per
// site.
__global__ __launch_bounds__(64) void k_el_nbr(ElecDev P, const double* __restrict__ x,
                                               const double* __restrict__ y, const double* __restrict__ z,
                                               int* __restrict__ nbrA, int* __restrict__ cntA,
                                               int* __restrict__ nbrP, int* __restrict__ cntP, int* __restrict__ err)
{
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    if (i >= P.n) return;
    const int wi = water_of(P, i);
    const bool poli = site_polr2(P, i) > 1.e-6;
    const double rl2 = P.rcut * P.rcut * (1.0 + 1e-9);
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int nA = 0, nP = 0;
    // candidates j (or -1) one per lane, kept in lane order
    auto take = [&](int j) {
        bool inA = false, inP = false;
        if (j >= 0 && water_of(P, j) != wi) {
            double d[3];
            const double r2 = sep(P, x, y, z, i, j, d);
            inA = r2 < rl2;
            inP = inA && poli && site_polr2(P, j) > 1.e-6;
        }
        const unsigned long long bA = __ballot(inA), bP = __ballot(inP);
        if (inA) {
            const int p = nA + __popcll(bA & below);
            if (p < P.capA) nbrA[(long)i * P.capA + p] = j;
        }
        if (inP) {
            const int p = nP + __popcll(bP & below);
            if (p < P.capP) nbrP[(long)i * P.capP + p] = j;
        }
        nA += __popcll(bA);
        nP += __popcll(bP);
    };
    if (P.cells.on) {
        // the 27 link cells around site i: cell order fixed, ascending index within a cell
        for_neighbour_cells(P.cells, i, [&](int c0, int c1) {
            for (int base = c0; base < c1; base += 64) take(base + lane < c1 ? P.cells.idx[base + lane] : -1);
        });
    } else {
        for (int base = 0; base < P.n; base += 64) take(base + lane < P.n ? base + lane : -1);
    }
    if (lane == 0) {
        cntA[i] = min(nA, P.capA);
        cntP[i] = min(nP, P.capP);
        if (nA > P.capA || nP > P.capP) atomicOr(err, 1);
    }
}

// Per-row set-up of the real-space fields (rfield.f:62-323): the charge field of the row's
// partners (iloop = 1 part, with the lads smearing) and the Thole-damped dipole tensors of
// its polarizable partners (stored SoA [6][n][capP]).  One wavefront per site.
__global__ __launch_bounds__(64) void k_el_prep(ElecDev P, const double* __restrict__ x,
                                                const double* __restrict__ y, const double* __restrict__ z,
                                                const double* __restrict__ chge, const double* __restrict__ tab,
                                                const int* __restrict__ nbrA, const int* __restrict__ cntA,
                                                const int* __restrict__ nbrP, const int* __restrict__ cntP,
                                                double* __restrict__ tens, double* __restrict__ efk)
{
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    if (i >= P.n) return;
    const double pi_ = site_polr(P, i);
    double e[3] = {0.0, 0.0, 0.0};
    const int na = cntA[i];
    for (int s = lane; s < na; s += 64) {
        const int j = nbrA[(long)i * P.capA + s];
        double d[3];
        const double r2 = sep(P, x, y, z, i, j, d);
        const double rr = sqrt(r2);
        if (!(P.rcut > rr)) continue;
        double bbb[4];
        table_interp(tab, P.drewd, rr, bbb);
        const double r3i = 1.0 / (rr * rr * rr);
        const double bb = pow(pi_ * site_polr(P, j), 1.0 / 6.0);
        const double u = rr / bb;
        const double arob4 = P.ascd * (u * u) * (u * u);
        const double dum1 = exp(-arob4);
        double eki = R4PIE0 * chge[j] * (bbb[1] + r3i);
        eki = eki - R4PIE0 * chge[j] * dum1 * r3i;
        e[0] += eki * d[0];
        e[1] += eki * d[1];
        e[2] += eki * d[2];
    }
    for (int q = 0; q < 3; ++q) e[q] = wsum(e[q]);
    if (lane == 0)
        for (int q = 0; q < 3; ++q) efk[3 * i + q] = e[q];
    const int np = cntP[i];
    const long nc = (long)P.n * P.capP;
    for (int s = lane; s < np; s += 64) {
        const int j = nbrP[(long)i * P.capP + s];
        double d[3];
        const double r2 = sep(P, x, y, z, i, j, d);
        const double rr = sqrt(r2);
        double dt[6] = {0, 0, 0, 0, 0, 0};
        if (P.rcut > rr) {
            double bbb[4];
            table_interp(tab, P.drewd, rr, bbb);
            const double rsqi = 1.0 / r2;
            const double r3i = 1.0 / (rr * rr * rr);
            const double r5i = r3i * rsqi;
            const double bb = pow(pi_ * site_polr(P, j), 1.0 / 6.0);
            const double u = rr / bb;
            const double arob4 = P.athole * (u * u) * (u * u);
            const double dum1 = exp(-arob4);
            const double dum2 = (1.0 + (4.0 / 3.0) * arob4) * dum1;
            const double rij[6] = {d[0] * d[0], d[0] * d[1], d[0] * d[2], d[1] * d[1], d[1] * d[2], d[2] * d[2]};
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const bool dg_ = (q == 0 || q == 3 || q == 5);
                dt[q] = R4PIE0 * ((3.0 * rij[q] * r5i - (dg_ ? r3i : 0.0)) + rij[q] * bbb[2] - (dg_ ? bbb[1] : 0.0));
                dt[q] = dt[q] - (3.0 * rij[q] * r5i * dum2 - (dg_ ? dum1 * r3i : 0.0)) * R4PIE0;
            }
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) tens[q * nc + (long)i * P.capP + s] = dt[q];
    }
}

// Excluded (intramolecular) pairs of each water (rfield.f:327-525): charge field with the
// -erf kernel and the dipole tensors (Ewald + Applequist - Thole with athole12/athole13),
// stored [nw][pair][6] for pairs (O,H1) (O,H2) (O,M) (H1,H2) (H1,M) (H2,M).
__global__ void k_el_intra_prep(ElecDev P, const double* __restrict__ x, const double* __restrict__ y,
                                const double* __restrict__ z, const double* __restrict__ chge,
                                double* __restrict__ tin, double* __restrict__ efk)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= P.nw) return;
    const int s[4] = {3 * w, 3 * w + 1, 3 * w + 2, P.nreal + w};
    int pr = 0;
    for (int a = 0; a < 4; ++a)
        for (int b = a + 1; b < 4; ++b, ++pr) {
            const int i = s[a], j = s[b];
            double d[3];
            const double r2 = sep(P, x, y, z, i, j, d);
            const double rr = sqrt(r2);
            const double rsqi = 1.0 / r2;
            const double r3i = 1.0 / (rr * r2);
            const double r5i = r3i * rsqi;
            double bbb[4];
            bfun_direct(P.alpha, rr, rsqi, bbb);
            const double eki = R4PIE0 * chge[j] * bbb[1];
            const double ekj = R4PIE0 * chge[i] * bbb[1];
            for (int q = 0; q < 3; ++q) {
                efk[3 * i + q] += eki * d[q];
                efk[3 * j + q] -= ekj * d[q];
            }
            double dt[6] = {0, 0, 0, 0, 0, 0};
            if (site_polr2(P, i) > 1.e-6 && site_polr2(P, j) > 1.e-6) {
                const double bb = pow(site_polr(P, i) * site_polr(P, j), 1.0 / 6.0);
                const double u = rr / bb;
                const double ath = (i % 3 == 0 || j % 3 == 0) ? P.athole12 : P.athole13;
                const double arob4 = ath * (u * u) * (u * u);
                const double dum1 = exp(-arob4);
                const double dum2 = (1.0 + (4.0 / 3.0) * arob4) * dum1;
                const double rij[6] = {d[0] * d[0], d[0] * d[1], d[0] * d[2], d[1] * d[1], d[1] * d[2], d[2] * d[2]};
                for (int q = 0; q < 6; ++q) {
                    const bool dg_ = (q == 0 || q == 3 || q == 5);
                    dt[q] = rij[q] * bbb[2] - (dg_ ? bbb[1] : 0.0);
                    dt[q] = dt[q] + 3.0 * rij[q] * r5i - (dg_ ? r3i : 0.0);
                    dt[q] = dt[q] - 3.0 * rij[q] * r5i * dum2 + (dg_ ? dum1 * r3i : 0.0);
                }
            }
            for (int q = 0; q < 6; ++q) tin[36 * w + 6 * pr + q] = dt[q];
        }
}

// Reciprocal-space phase tables of each site: cos / sin (l * 2 pi * s_a), s = rcell . r,
// l = 0..kmax_a, by the reference's recurrence (ewald1p.f:196-241, 262-282), layout
// phc/phs [a][l][n].
__global__ void k_el_phase(ElecDev P, const double* __restrict__ x, const double* __restrict__ y,
                           const double* __restrict__ z, const double* __restrict__ rcell,
                           double* __restrict__ phc, double* __restrict__ phs)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P.n) return;
    const int ks = P.kstride;
    for (int a = 0; a < 3; ++a) {
        const double ss = rcell[a] * x[j] + rcell[3 + a] * y[j] + rcell[6 + a] * z[j];
        const double c1 = cos(TWOPI * ss), s1 = sin(TWOPI * ss);
        double c = 1.0, s = 0.0;
        for (int l = 0; l <= P.km[a]; ++l) {
            if (l == 1) {
                c = c1;
                s = s1;
            } else if (l > 1) {
                const double cs = c;
                c = cs * c1 - s * s1;
                s = s * c1 + cs * s1;
            }
            phc[((long)a * ks + l) * P.n + j] = c;
            phs[((long)a * ks + l) * P.n + j] = s;
        }
    }
}

// ======================================================= reciprocal-space passes ======

namespace {

// exp(i k.r_j) at the first k of a run (l, m, n0), with the reference's sign rules
// (ewald1p.f:290-306 for m, 346-352 / 392-398 for n).
__device__ __forceinline__ void run_phase(double el_c, double el_s, double em_c, double em_s, double en_c, double en_s,
                                          int m, int n0, double& ckr, double& skr)
{
    // negative m or n use the conjugate table entry; the selects keep the wavefront
    // convergent and give the same products as the reference's branches
    const double sm = m >= 0 ? em_s : -em_s;
    const double sn = n0 >= 0 ? en_s : -en_s;
    const double clm = el_c * em_c - el_s * sm;
    const double slm = el_s * em_c + sm * el_c;
    ckr = clm * en_c - slm * sn;
    skr = slm * en_c + clm * sn;
}

}  // namespace

// Structure factors over a chunk of sites, k-parallel (ewald1p.f:342-438).  A block owns
// 64 half-runs (lane = half-run: up to EL_HALF consecutive k at fixed l, m, stepping
// exp(i k.r_j) along n by a complex recurrence) and EL_SG site groups (one wavefront
// each): the block stages the phase tables of its site chunk in LDS, EL_TILE sites at a
// time (rows l = lmin..lmax of a*, every |m| row of b*, every |n| row of c*, padded against
// bank conflicts), wavefront w takes the tile's sites w, w + EL_SG, ..., and the EL_SG
// partial sums of every k are added in LDS in group order before one partial per chunk is
// written: Spart[chunk][k][NV], NV = 8 with charges (sum q ckr, sum q skr, sum mu ckr (3),
// sum mu skr (3)), NV = 6 for the dipoles alone (the last six).  The dipole sums run over
// polarizable sites only.
#ifndef EL_TILE
#define EL_TILE 16
#endif
#define EL_TROW (EL_TILE + 1)
#ifndef EL_SPLIT
#define EL_SPLIT 2                    // threads per k run in the structure-factor pass
#endif
#ifndef EL_SG
#define EL_SG 4                       // site groups (wavefronts) per structure-factor block
#endif
#define EL_HALF (EL_RUN / EL_SPLIT)
// With cgblk (the conjugate-gradient solver's partials, cgnb blocks), the pass first takes the
// solver's stopping test for iteration `it` (cg_step; block (0, 0) records the scalars) and
// does nothing once the dipoles have converged.
template <bool CHG>
__global__ __launch_bounds__(64 * EL_SG) void k_el_sfac(ElecDev P, int spc, const int* __restrict__ runs,
                                                        const int* __restrict__ runk, const double* __restrict__ phc,
                                                        const double* __restrict__ phs, const double* __restrict__ chge,
                                                        const double* __restrict__ dip, double* __restrict__ Spart,
                                                        double* __restrict__ state, const double* __restrict__ cgblk,
                                                        int cgnb, int it)
{
    if (!CHG && state[2] != 0.0) return;
    if (!CHG && cgblk &&
        cg_step(P, it, cgnb, cgblk, state, blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 64))
        return;
    extern __shared__ double lds[];
    constexpr int NT = 64 * EL_SG;
    constexpr int NV = CHG ? 8 : 6;
    constexpr int Q0 = CHG ? 0 : 2;          // first accumulator component written
    const int tid = threadIdx.x, lane = tid & 63, grp = tid >> 6;
    const int g = blockIdx.x * 64 + lane;
    const int r = g / EL_SPLIT, half = g % EL_SPLIT;
    const int chunk = blockIdx.y;
    const int ns = CHG ? P.n : P.nrow;
    const int j0 = chunk * spc, j1 = min(ns, j0 + spc);
    const int rfirst = (blockIdx.x * 64) / EL_SPLIT;
    const int rlast = min(P.nrun - 1, (blockIdx.x * 64 + 63) / EL_SPLIT);
    const int lmin = runs[4 * rfirst], lmax = runs[4 * rlast];
    const int nl = lmax - lmin + 1, nm = P.km[1] + 1, nn = P.km[2] + 1;
    const int nrows = nl + nm + nn;
    // (cos, sin) pairs of the tile's table rows, and (q, mu) of its sites, as double2 so
    // that each entry is one 16-byte LDS read (dynamic LDS starts 16-byte aligned)
    double2* tcs = reinterpret_cast<double2*>(lds);
    double2* tw = tcs + nrows * EL_TROW;
    const int ks = P.kstride;
    int l = lmin, m = 0, nst = 1, cnt = 0;
    if (r < P.nrun) {
        l = runs[4 * r];
        m = runs[4 * r + 1];
        nst = runs[4 * r + 2] + EL_HALF * half;
        cnt = min(EL_HALF, max(0, runs[4 * r + 3] - EL_HALF * half));
        if (cnt == 0) nst = runs[4 * r + 2];
    }
    const int rowA = (l - lmin) * EL_TROW, rowB = (nl + abs(m)) * EL_TROW, rowC = (nl + nm + abs(nst)) * EL_TROW;
    const int rowC1 = (nl + nm + 1) * EL_TROW;
    double acc[EL_HALF][8];
#pragma unroll
    for (int t = 0; t < EL_HALF; ++t)
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[t][q] = 0.0;
    // tiles of EL_TILE sites, double-buffered: the next tile's table entries are loaded
    // into registers while the current tile is consumed from LDS
    constexpr int PF = (80 * EL_TILE + NT - 1) / NT;   // staged entries per thread (nrows <= 80)
    double pc[PF], ps[PF], pw = 0.0;
    auto fetch = [&](int t0) {
        const int tn = min(EL_TILE, j1 - t0);
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int idx = tid + NT * u;
            const int row = idx / EL_TILE, sidx = idx - row * EL_TILE;
            pc[u] = 0.0;
            ps[u] = 0.0;
            if (row < nrows && sidx < tn) {
                int a, ll;
                if (row < nl) {
                    a = 0;
                    ll = lmin + row;
                } else if (row < nl + nm) {
                    a = 1;
                    ll = row - nl;
                } else {
                    a = 2;
                    ll = row - nl - nm;
                }
                const long o = ((long)a * ks + ll) * P.n + t0 + sidx;
                pc[u] = phc[o];
                ps[u] = phs[o];
            }
        }
        if (tid < 4 * EL_TILE) {
            // a site without an induced dipole carries mu = 0, so no test in the sums
            const int sj = tid >> 2, w = tid & 3;
            const int j = t0 + sj;
            pw = 0.0;
            if (sj < tn) {
                if (w == 0) pw = chge[j];
                else if (site_polr2(P, j) > 1.e-6) pw = dip[3 * j + w - 1];
            }
        }
    };
    if (j0 < j1) fetch(j0);
    for (int t0 = j0; t0 < j1; t0 += EL_TILE) {
        const int tn = min(EL_TILE, j1 - t0);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int idx = tid + NT * u;
            const int row = idx / EL_TILE, sidx = idx - row * EL_TILE;
            if (row < nrows) tcs[row * EL_TROW + sidx] = make_double2(pc[u], ps[u]);
        }
        if (tid < 4 * EL_TILE) reinterpret_cast<double*>(tw)[tid] = pw;
        __syncthreads();
        if (t0 + EL_TILE < j1) fetch(t0 + EL_TILE);
        if (cnt > 0) {
            // every thread steps EL_HALF k; terms past the end of its run are never stored
#pragma unroll 2
            for (int sj = grp; sj < tn; sj += EL_SG) {
                const double2 A = tcs[rowA + sj], B = tcs[rowB + sj], C = tcs[rowC + sj], C1 = tcs[rowC1 + sj];
                const double2 w01 = tw[2 * sj], w23 = tw[2 * sj + 1];
                double ckr, skr;
                run_phase(A.x, A.y, B.x, B.y, C.x, C.y, m, nst, ckr, skr);
                const double q = w01.x, mx = w01.y, my = w23.x, mz = w23.y;
#pragma unroll
                for (int t = 0; t < EL_HALF; ++t) {
                    if (CHG) {
                        acc[t][0] += q * ckr;
                        acc[t][1] += q * skr;
                    }
                    acc[t][2] += ckr * mx;
                    acc[t][3] += ckr * my;
                    acc[t][4] += ckr * mz;
                    acc[t][5] += skr * mx;
                    acc[t][6] += skr * my;
                    acc[t][7] += skr * mz;
                    const double cn = ckr * C1.x - skr * C1.y;
                    skr = skr * C1.x + ckr * C1.y;
                    ckr = cn;
                }
            }
        }
    }
    // the site groups' sums, added in group order through LDS (the tiles are done with it);
    // every thread reaches every barrier
    double* red = lds;                                   // [EL_SG - 1][64][NV]
#pragma unroll
    for (int t = 0; t < EL_HALF; ++t) {
        __syncthreads();
        if (grp > 0)
#pragma unroll
            for (int q = 0; q < NV; ++q) red[((grp - 1) * 64 + lane) * NV + q] = acc[t][Q0 + q];
        __syncthreads();
        if (grp == 0)
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                double v = acc[t][Q0 + q];
                for (int w = 1; w < EL_SG; ++w) v += red[((w - 1) * 64 + lane) * NV + q];
                acc[t][Q0 + q] = v;
            }
    }
    if (grp != 0 || cnt == 0) return;
    const int k0 = runk[r] + EL_HALF * half;
#pragma unroll
    for (int t = 0; t < EL_HALF; ++t)
        if (t < cnt)
#pragma unroll
            for (int q = 0; q < NV; ++q) Spart[((long)chunk * P.K + k0 + t) * NV + q] = acc[t][Q0 + q];
}

// Sum the chunk partials and derive the per-k constants of the field passes: kc[k][0..5] =
// (-k_a akk k.Ms, -k_a akk k.Mc) for the dipole field (ewald1p.f:759-767).  With CHG the
// charge sums are stored in Sq[k][2] and kc[k][6..11] receives the charge-field constants
// (k_a akk ckcs, -k_a akk ckss) (ewald1p.f:749-757).  A block sums 64 (k, component)
// elements (8 k vectors); lanes 8k'..8k'+7 of the first wave hold the 8 sums of one k.
template <bool CHG>
__device__ void sreduce_block(const ElecDev& P, int blk, int schunks, const double* __restrict__ Spart,
                              const double* __restrict__ kv, double* __restrict__ S, double* __restrict__ Sq,
                              double* __restrict__ kc)
{
    // NV values per k in the partials (see k_el_sfac); a block sums EPB consecutive
    // elements, a whole number of k
    constexpr int NV = CHG ? 8 : 6, EPB = CHG ? 64 : 60, Q0 = CHG ? 0 : 2;
    const int lane = threadIdx.x & 63;
    const long e = (long)blk * EPB + lane;
    const double v = chunk_sum64(Spart, (long)NV * P.K, schunks, e, lane < EPB && e < (long)NV * P.K);
    if (threadIdx.x >= 64) return;
    const int base = lane - lane % NV;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < NV; ++q) s[Q0 + q] = __shfl(v, min(base + q, 63), 64);
    const int k = (int)(e / NV);
    if (lane % NV != 0 || lane >= EPB || k >= P.K) return;
    if (CHG) {
        Sq[2 * k] = s[0];
        Sq[2 * k + 1] = s[1];
    }
#pragma unroll
    for (int q = 2; q < 8; ++q) S[8 * k + q] = s[q];
    const double kx = kv[8 * k], ky = kv[8 * k + 1], kz = kv[8 * k + 2], akk = kv[8 * k + 3];
    const double skdmuckrs = kx * s[2] + ky * s[3] + kz * s[4];
    const double skdmuskrs = kx * s[5] + ky * s[6] + kz * s[7];
    double* c = kc + 16 * k;
    c[0] = -kx * akk * skdmuskrs;
    c[1] = -ky * akk * skdmuskrs;
    c[2] = -kz * akk * skdmuskrs;
    c[3] = -kx * akk * skdmuckrs;
    c[4] = -ky * akk * skdmuckrs;
    c[5] = -kz * akk * skdmuckrs;
    if (CHG) {
        c[6] = kx * akk * s[0];
        c[7] = ky * akk * s[0];
        c[8] = kz * akk * s[0];
        c[9] = -kx * akk * s[1];
        c[10] = -ky * akk * s[1];
        c[11] = -kz * akk * s[1];
    }
}

template <bool CHG>
__global__ __launch_bounds__(256) void k_el_sreduce(ElecDev P, int schunks, const double* __restrict__ Spart,
                                                    const double* __restrict__ kv, double* __restrict__ S,
                                                    double* __restrict__ Sq, double* __restrict__ kc,
                                                    const double* __restrict__ state)
{
    if (!CHG && state[2] != 0.0) return;
    sreduce_block<CHG>(P, blockIdx.x, schunks, Spart, kv, S, Sq, kc);
}

// Reciprocal-space fields over a chunk of runs, site-parallel (ewald1p.f:749-769):
// efdd_a += c_a skr + d_a ckr (dipole field), with CHG also efdc_a += e_a skr + f_a ckr
// (charge field).  A block covers 128 sites (two per lane: j and j + 64) and EL_KG run
// groups (one wavefront each, runs g, g + EL_KG, ... of the chunk); it stages the chunk's
// runs and per-k constants in LDS, and the run groups' sums are added in LDS in group
// order before one partial per chunk is written: Fpart[chunk][6][n] (components 0..2 the
// dipole field, 3..5 the charge field).
#ifndef EL_KG
#define EL_KG 4                       // run groups (wavefronts) per field block
#endif
template <bool CHG>
__global__ __launch_bounds__(64 * EL_KG) void k_el_kfield(ElecDev P, const int* __restrict__ runs,
                                                          const int* __restrict__ runk, const double* __restrict__ phc,
                                                          const double* __restrict__ phs, const double* __restrict__ kc,
                                                          double* __restrict__ Fpart, const double* __restrict__ state)
{
    if (!CHG && state[2] != 0.0) return;
    extern __shared__ double lds[];
    constexpr int NC = CHG ? 12 : 6;
    constexpr int NT = 64 * EL_KG;
    const int tid = threadIdx.x, lane = tid & 63, grp = tid >> 6;
    const int chunk = blockIdx.y;
    const int r0 = chunk * P.frpc, r1 = min(P.nrun, r0 + P.frpc);
    const int nr = max(0, r1 - r0);
    // two sites per lane share every constant read; lanes past the last site compute on a
    // clamped site and store nothing (every thread reaches the barriers below).  The n = 1
    // phase factors are loaded before the staging barrier, which completes them: loaded
    // later, the wait for them inside the k loop would also wait for the run prefetches.
    const int nrow = CHG ? P.n : P.nrow;
    const int ja = blockIdx.x * 128 + lane, jb = ja + 64;
    const bool ha = ja < nrow, hb = jb < nrow;
    const int jA = ha ? ja : 0, jB = hb ? jb : jA;
    const int ks = P.kstride;
    const long o1 = ((long)2 * ks + 1) * P.n;
    const double ca1 = phc[o1 + jA], sa1 = phs[o1 + jA], cb1 = phc[o1 + jB], sb1 = phs[o1 + jB];
    double* sk = lds;                                          // [fkmax][NC]
    int* sr = reinterpret_cast<int*>(lds + (long)P.fkmax * NC); // [frpc][4] runs
    int* so = sr + 4 * P.frpc;                                  // [frpc] first k of each run - kbase
    const int kbase = nr ? runk[r0] : 0;
    const int nk = nr ? ((r1 < P.nrun ? runk[r1] : P.K) - kbase) : 0;
    for (int idx = tid; idx < nk * NC; idx += NT) {
        const int kk = idx / NC, q = idx - kk * NC;
        sk[idx] = kc[16L * (kbase + kk) + q];
    }
    for (int idx = tid; idx < nr * 4; idx += NT) sr[idx] = runs[4 * r0 + idx];
    for (int idx = tid; idx < nr; idx += NT) so[idx] = runk[r0 + idx] - kbase;
    __syncthreads();
    double fa[3] = {0, 0, 0}, fb[3] = {0, 0, 0}, ga[3] = {0, 0, 0}, gb[3] = {0, 0, 0};
    // the phase-table entries of a run are loaded one run ahead (double-buffered in
    // registers), so their latency overlaps the previous run's arithmetic
    double qa[6], qb[6];
    auto load_run = [&](int r) {
        const int l = sr[4 * r], m = sr[4 * r + 1], n0 = sr[4 * r + 2];
        const long oa = (long)l * P.n, ob = ((long)ks + abs(m)) * P.n, oc = ((long)2 * ks + abs(n0)) * P.n;
        qa[0] = phc[oa + jA]; qa[1] = phs[oa + jA]; qa[2] = phc[ob + jA];
        qa[3] = phs[ob + jA]; qa[4] = phc[oc + jA]; qa[5] = phs[oc + jA];
        qb[0] = phc[oa + jB]; qb[1] = phs[oa + jB]; qb[2] = phc[ob + jB];
        qb[3] = phs[ob + jB]; qb[4] = phc[oc + jB]; qb[5] = phs[oc + jB];
    };
    if (grp < nr) load_run(grp);
    for (int r = grp; r < nr; r += EL_KG) {
        const int m = sr[4 * r + 1], n0 = sr[4 * r + 2], cnt = sr[4 * r + 3], kk = so[r];
        double cA, sA, cB, sB;
        run_phase(qa[0], qa[1], qa[2], qa[3], qa[4], qa[5], m, n0, cA, sA);
        run_phase(qb[0], qb[1], qb[2], qb[3], qb[4], qb[5], m, n0, cB, sB);
        if (r + EL_KG < nr) load_run(r + EL_KG);
#pragma unroll
        for (int t = 0; t < EL_RUN; ++t) {
            if (t < cnt) {
                // one fused multiply-add per product into the accumulators
                const double* c = sk + NC * (kk + t);
                const double c0 = c[0], c1v = c[1], c2 = c[2], c3 = c[3], c4 = c[4], c5 = c[5];
                fa[0] += c0 * sA;
                fa[0] += c3 * cA;
                fa[1] += c1v * sA;
                fa[1] += c4 * cA;
                fa[2] += c2 * sA;
                fa[2] += c5 * cA;
                fb[0] += c0 * sB;
                fb[0] += c3 * cB;
                fb[1] += c1v * sB;
                fb[1] += c4 * cB;
                fb[2] += c2 * sB;
                fb[2] += c5 * cB;
                if (CHG) {
                    const double c6 = c[6], c7 = c[7], c8 = c[8], c9 = c[9], c10 = c[10], c11 = c[11];
                    ga[0] += c6 * sA;
                    ga[0] += c9 * cA;
                    ga[1] += c7 * sA;
                    ga[1] += c10 * cA;
                    ga[2] += c8 * sA;
                    ga[2] += c11 * cA;
                    gb[0] += c6 * sB;
                    gb[0] += c9 * cB;
                    gb[1] += c7 * sB;
                    gb[1] += c10 * cB;
                    gb[2] += c8 * sB;
                    gb[2] += c11 * cB;
                }
                const double na = cA * ca1 - sA * sa1;
                sA = sA * ca1 + cA * sa1;
                cA = na;
                const double nb = cB * cb1 - sB * sb1;
                sB = sB * cb1 + cB * sb1;
                cB = nb;
            }
        }
    }
    // the run groups' sums, added in group order through LDS
    constexpr int NQ = CHG ? 12 : 6;                           // values per lane
    double v[12];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        v[q] = fa[q];
        v[3 + q] = fb[q];
        if (CHG) {
            v[6 + q] = ga[q];
            v[9 + q] = gb[q];
        }
    }
    double* red = lds;                                         // [EL_KG - 1][64][NQ]
    __syncthreads();
    if (grp > 0)
#pragma unroll
        for (int q = 0; q < NQ; ++q) red[((grp - 1) * 64 + lane) * NQ + q] = v[q];
    __syncthreads();
    if (grp != 0) return;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        for (int w = 1; w < EL_KG; ++w) v[q] += red[((w - 1) * 64 + lane) * NQ + q];
    double* out = Fpart + (long)chunk * 6 * P.n;
    for (int q = 0; q < 3; ++q) {
        if (ha) out[(long)q * P.n + ja] = v[q];
        if (hb) out[(long)q * P.n + jb] = v[3 + q];
        if (CHG) {
            if (ha) out[(long)(3 + q) * P.n + ja] = v[6 + q];
            if (hb) out[(long)(3 + q) * P.n + jb] = v[9 + q];
        }
    }
}

// Sum the field partials over the k chunks and scale (ewald1p.f:941-976): efdd =
// scale * sum + self term; with CHG also efdc = scale * sum (iloop = 1).  Element
// (component q, site j) = q * nrow + j, 64 per block.
template <bool CHG>
__global__ __launch_bounds__(256) void k_el_kfield_reduce(ElecDev P, const double* __restrict__ Fpart,
                                                          const double* __restrict__ dip, double* __restrict__ efdd,
                                                          double* __restrict__ efdc, const double* __restrict__ state)
{
    if (!CHG && state[2] != 0.0) return;
    const int nrow = CHG ? P.n : P.nrow;
    const int nq = CHG ? 6 : 3;
    const int t = blockIdx.x * 64 + (threadIdx.x & 63);
    const bool valid = t < nq * nrow;
    const int q = valid ? t / nrow : 0, j = valid ? t - q * nrow : 0;
    const double sum = chunk_sum64(Fpart, 6L * P.n, P.fchunks, (long)q * P.n + j, valid);
    if (threadIdx.x >= 64 || !valid) return;
    if (q < 3) {
        const bool pol = site_polr2(P, j) > 1.e-6;
        efdd[3 * j + q] = (pol ? P.scale * sum : 0.0) + P.self_dd * dip[3 * j + q];
    } else {
        efdc[3 * j + q - 3] = P.scale * sum;
    }
}

// ============================================================ solver iteration ======

// Real-space dipole field of one row (rfield.f:201-317 list pairs; 419-520 excluded
// pairs): emu_i = sum_j T_ij mu_j.  One wavefront per site.
__device__ void emu_row(const ElecDev& P, int i, int lane, const int* __restrict__ nbrP, const int* __restrict__ cntP,
                        const double* __restrict__ tens, const double* __restrict__ tin,
                        const double* __restrict__ dip, double* __restrict__ emu)
{
    if (i >= P.nrow) return;
    const long nc = (long)P.n * P.capP;
    double e[3] = {0, 0, 0};
    const int np = cntP[i];
    for (int s = lane; s < np; s += 64) {
        const long o = (long)i * P.capP + s;
        const int j = nbrP[o];
        const double mx = dip[3 * j], my = dip[3 * j + 1], mz = dip[3 * j + 2];
        const double t0 = tens[o], t1 = tens[nc + o], t2 = tens[2 * nc + o];
        const double t3 = tens[3 * nc + o], t4 = tens[4 * nc + o], t5 = tens[5 * nc + o];
        e[0] += t0 * mx + t1 * my + t2 * mz;
        e[1] += t1 * mx + t3 * my + t4 * mz;
        e[2] += t2 * mx + t4 * my + t5 * mz;
    }
    // excluded partners of the same water on lanes 0 and 1: the other two real sites (the M
    // site carries no induced dipole and its entries are not part of the solve); pair order
    // of tin: (O,H1) (O,H2) (O,M) (H1,H2) (H1,M) (H2,M)
    if (lane < 2) {
        const int w = i / 3, si = i - 3 * w;
        const int pt = lane < si ? lane : lane + 1;
        const int a = min(si, pt), b = max(si, pt);
        const int pr = a == 0 ? b - 1 : 3;
        const int j = 3 * w + pt;
        if (site_polr2(P, j) > 1.e-6) {
            const double* dt = tin + 36 * w + 6 * pr;
            const double mx = dip[3 * j] * R4PIE0, my = dip[3 * j + 1] * R4PIE0, mz = dip[3 * j + 2] * R4PIE0;
            e[0] += dt[0] * mx + dt[1] * my + dt[2] * mz;
            e[1] += dt[1] * mx + dt[3] * my + dt[4] * mz;
            e[2] += dt[2] * mx + dt[4] * my + dt[5] * mz;
        }
    }
    for (int q = 0; q < 3; ++q) e[q] = wsum(e[q]);
    if (lane == 0)
        for (int q = 0; q < 3; ++q) emu[3 * i + q] = e[q];
}

// emu for four rows per block.  With cgblk (the conjugate-gradient solver's partials, cgnb
// blocks), the solver's stopping test for iteration `it` comes first (cg_step; block 0
// records the scalars) and the kernel does nothing once the dipoles have converged.
__global__ __launch_bounds__(256) void k_el_emu(ElecDev P, const int* __restrict__ nbrP, const int* __restrict__ cntP,
                                                const double* __restrict__ tens, const double* __restrict__ tin,
                                                const double* __restrict__ dip, double* __restrict__ emu,
                                                double* __restrict__ state, const double* __restrict__ cgblk, int cgnb,
                                                int it)
{
    if (state[2] != 0.0) return;
    if (cgblk && cg_step(P, it, cgnb, cgblk, state, blockIdx.x == 0 && threadIdx.x < 64)) return;
    emu_row(P, blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63, nbrP, cntP, tens, tin, dip, emu);
}

// One launch for two independent parts of an application of the dipole-field operator:
// blocks [0, nsr) sum the dipole structure factors (k_el_sreduce<false>), the others form
// the real-space field emu, four rows per block.
__global__ __launch_bounds__(256) void k_el_sreduce_emu(ElecDev P, int nsr, int schunks, const double* __restrict__ Spart,
                                                        const double* __restrict__ kv, double* __restrict__ S,
                                                        double* __restrict__ Sq, double* __restrict__ kc,
                                                        const int* __restrict__ nbrP, const int* __restrict__ cntP,
                                                        const double* __restrict__ tens, const double* __restrict__ tin,
                                                        const double* __restrict__ dip, double* __restrict__ emu,
                                                        const double* __restrict__ state)
{
    if (state[2] != 0.0) return;
    const int b = blockIdx.x;
    if (b < nsr) {
        sreduce_block<false>(P, b, schunks, Spart, kv, S, Sq, kc);
        return;
    }
    emu_row(P, (b - nsr) * 4 + (threadIdx.x >> 6), threadIdx.x & 63, nbrP, cntP, tens, tin, dip, emu);
}

// SOR update (iter.f:58-89) and per-block sums of |dmu|^2 and of the dipole count.
__global__ __launch_bounds__(256) void k_el_sor(ElecDev P, const double* __restrict__ efk,
                                                const double* __restrict__ emu, const double* __restrict__ efdd,
                                                const double* __restrict__ efdc, double* __restrict__ dip,
                                                double* __restrict__ blk, const double* __restrict__ state)
{
    if (state[2] != 0.0) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < P.nrow) {
        const double p2 = site_polr2(P, i);
        if (p2 > P.toler) {
            const double oldfac = P.omega, fac = 1.0 - oldfac;
            double dd = 0.0;
            for (int q = 0; q < 3; ++q) {
                const double old = dip[3 * i + q];
                const double nw = oldfac * old + fac * p2 *
                                                      (efk[3 * i + q] + emu[3 * i + q] + efdd[3 * i + q] +
                                                       efdc[3 * i + q] + 0.0) /
                                                      R4PIE0;
                dip[3 * i + q] = nw;
                dd += (nw - old) * (nw - old);
            }
            v[0] = dd;
            v[1] = 1.0;
        }
    }
    block_sum_store(v, blk + 2 * blockIdx.x);
}

// Convergence test (iter.f:95-101): state = [iterations, rms dmu, done, ...].
__global__ __launch_bounds__(64) void k_el_converge(ElecDev P, int nb, const double* __restrict__ blk,
                                                    double* __restrict__ state)
{
    if (state[2] != 0.0) return;
    double s = 0.0, c = 0.0;
    for (int b = threadIdx.x; b < nb; b += 64) {
        s += blk[2 * b];
        c += blk[2 * b + 1];
    }
    s = wsum(s);
    c = wsum(c);
    if (threadIdx.x == 0) {
        if (c == 0.0) c = 1.0;
        const double delta = sqrt(s / c);
        state[0] += 1.0;
        state[1] = delta;
        if (delta <= P.toler) state[2] = 1.0;
    }
}


// ==================================================== conjugate-gradient solver ======
// The induced dipoles solve A mu = E0 with A = D - T, D = r4pie0 / polr2 (per site), T the
// dipole-field operator (emu + efdd) and E0 = efk + efdc.  A is symmetric positive definite
// for Thole-damped dipoles, so preconditioned conjugate gradients reach the SOR fixed point
// of iter.f in far fewer iterations.  The preconditioner M is block-Jacobi over waters: the
// 9 x 9 block of each water's O, H1, H2 dipoles, which holds the strong intramolecular
// couplings (half the iterations of a diagonal preconditioner at liquid density).  The
// stopping test is the reference's: the RMS of the step an SOR update would take from the
// current dipoles, (1 - omega) * rms(polr2 r / r4pie0), against toler (iter.f:95-101).
//
// The iteration is the Chronopoulos-Gear form of preconditioned conjugate gradients
// (A. T. Chronopoulos and C. W. Gear, J. Comput. Appl. Math. 25 (1989) 153-168), in which
// both inner products of an iteration are taken after its vector update, so that one
// iteration is one application of A and one fused update with a single set of block
// partials:
//   n = A m
//   z = n + beta z,  q = m + beta q,  s = w + beta s,  p = u + beta p
//   mu += alpha p,  r -= alpha s,  u -= alpha q,  w -= alpha z,  m = M^-1 w
//   partials of g = r.u and d = w.u, from which cg_step forms the next alpha and beta
// (u = M^-1 r, w = A u, s = A p, q = M^-1 s, z = A q throughout).  It starts from
// r = E0 - A mu0, u = M^-1 r and one extra application of A for w = A u.  T mu, which the
// final passes need, follows mu as T mu += alpha (D p - s); with the Ewald sum over k the
// dipole structure factors follow it too: S(q) = S(m) + beta S(q), S(p) = S(u) + beta S(p),
// S(mu) += alpha S(p), S(u) -= alpha S(q).  The scalars need no grid-wide synchronization:
// the first kernel of each iteration sums the update's block partials in every wavefront
// (cg_step) and its first block records them for the update.

// S(mu) of the start dipoles (the dipole part of the preparation's structure factors)
__global__ __launch_bounds__(256) void k_cg_smu_init(ElecDev P, const double* __restrict__ S, double* __restrict__ Smu)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e < 6L * P.K) Smu[e] = S[8 * (e / 6) + 2 + e % 6];
}

// Fields and structure factors of the converged dipoles for the final passes, from the
// quantities the solver carried along with mu: emu = T mu (real and reciprocal parts
// together), efdd = 0, S = S(mu)
__global__ __launch_bounds__(256) void k_cg_final_fields(ElecDev P, const double* __restrict__ tmu,
                                                         const double* __restrict__ Smu, double* __restrict__ emu,
                                                         double* __restrict__ efdd, double* __restrict__ S)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e < 3L * P.nrow) {
        emu[e] = tmu[e];
        efdd[e] = 0.0;
    }
    if (e < 6L * P.K) S[8 * (e / 6) + 2 + e % 6] = Smu[e];
}

// packed index of (i, j) in a symmetric 9 x 9 matrix stored as its lower triangle
__device__ __forceinline__ constexpr int tri9(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

// Per water: the preconditioner block M = diag(r4pie0 / polr2) - r4pie0 T_damped over the
// intramolecular pairs (the Thole-damped dipole tensor of rfield.f:419-520 without its Ewald
// split: M only has to be close to A's block, it does not move the fixed point), inverted
// by Cholesky and stored packed (45 values per water).  Rows of sites without a dipole are
// decoupled unit rows.  A block that is not positive definite falls back to the diagonal.
__global__ __launch_bounds__(256) void k_cg_block(ElecDev P, const double* __restrict__ x,
                                                  const double* __restrict__ y, const double* __restrict__ z,
                                                  double* __restrict__ minv)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= P.nw) return;
    double a[45];
    bool on[3];
#pragma unroll
    for (int q = 0; q < 45; ++q) a[q] = 0.0;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const double p2 = site_polr2(P, 3 * w + s);
        on[s] = p2 > P.toler;
#pragma unroll
        for (int c = 0; c < 3; ++c) a[tri9(3 * s + c, 3 * s + c)] = on[s] ? R4PIE0 / p2 : 1.0;
    }
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int t = s + 1; t < 3; ++t) {
            const int i = 3 * w + s, j = 3 * w + t;
            if (!(on[s] && on[t] && site_polr2(P, i) > 1.e-6 && site_polr2(P, j) > 1.e-6)) continue;
            double d[3];
            const double r2 = sep(P, x, y, z, i, j, d);
            const double rr = sqrt(r2);
            const double r3i = 1.0 / (rr * r2), r5i = r3i / r2;
            const double bb = pow(site_polr(P, i) * site_polr(P, j), 1.0 / 6.0);
            const double u = rr / bb;
            const double arob4 = (s == 0 ? P.athole12 : P.athole13) * (u * u) * (u * u);
            const double dum1 = exp(-arob4);
            const double dum2 = (1.0 + (4.0 / 3.0) * arob4) * dum1;
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int e = 0; e < 3; ++e)
                    a[tri9(3 * t + e, 3 * s + c)] =
                        -R4PIE0 * (3.0 * d[c] * d[e] * r5i * (1.0 - dum2) - (c == e ? r3i * (1.0 - dum1) : 0.0));
        }
    // Cholesky M = L L^T in place
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double v = a[tri9(i, j)];
#pragma unroll
            for (int k = 0; k < j; ++k) v -= a[tri9(i, k)] * a[tri9(j, k)];
            if (i == j) {
                ok = ok && v > 0.0;
                a[tri9(i, i)] = sqrt(fmax(v, 1.e-300));
            } else {
                a[tri9(i, j)] = v / a[tri9(j, j)];
            }
        }
    // X = L^-1 (lower), then M^-1 = X^T X
    double X[45];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        X[tri9(i, i)] = 1.0 / a[tri9(i, i)];
#pragma unroll
        for (int j = 0; j < i; ++j) {
            double v = 0.0;
#pragma unroll
            for (int k = j; k < i; ++k) v += a[tri9(i, k)] * X[tri9(k, j)];
            X[tri9(i, j)] = -v / a[tri9(i, i)];
        }
    }
    double* m = minv + 45L * w;
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double v = 0.0;
#pragma unroll
            for (int k = i; k < 9; ++k) v += X[tri9(k, i)] * X[tri9(k, j)];
            ok = ok && isfinite(v);
            a[tri9(i, j)] = v;
        }
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            const int s = i / 3;
            const double jac = (i == j && on[s]) ? site_polr2(P, 3 * w + s) / R4PIE0 : 0.0;
            m[tri9(i, j)] = ok ? a[tri9(i, j)] : jac;
        }
}

// z = M^-1 r for one water (rows of sites without a dipole stay zero)
__device__ __forceinline__ void cg_precondition(const double* __restrict__ m, const double (&r)[9],
                                                const bool (&on)[3], double (&z)[9])
{
    double mm[45];
#pragma unroll
    for (int q = 0; q < 45; ++q) mm[q] = m[q];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        double v = 0.0;
#pragma unroll
        for (int j = 0; j < 9; ++j) v += mm[tri9(i, j)] * r[j];
        z[i] = on[i / 3] ? v : 0.0;
    }
}

// r = E0 + T mu0 - D mu0 (T mu0 = emu + efdd of the start dipoles), u = M^-1 r, T mu = T mu0;
// the first application of A goes to m = u; z, q, s, p start at zero.  One thread per water.
__global__ __launch_bounds__(256) void k_cg_init(ElecDev P, const double* __restrict__ efk,
                                                 const double* __restrict__ efdc, const double* __restrict__ emu,
                                                 const double* __restrict__ efdd, const double* __restrict__ dip,
                                                 const double* __restrict__ minv, double* __restrict__ r,
                                                 double* __restrict__ u, double* __restrict__ m, double* __restrict__ z,
                                                 double* __restrict__ qv, double* __restrict__ sv,
                                                 double* __restrict__ pv, double* __restrict__ tmu)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= P.nw) return;
    double rr[9], uu[9];
    bool on[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const double p2 = site_polr2(P, 3 * w + s);
        on[s] = p2 > P.toler;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int o = 9 * w + 3 * s + c;
            tmu[o] = emu[o] + efdd[o];
            rr[3 * s + c] = on[s] ? (efk[o] + efdc[o] + emu[o] + efdd[o]) - R4PIE0 * dip[o] / p2 : 0.0;
        }
    }
    cg_precondition(minv + 45L * w, rr, on, uu);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int o = 9 * w + q;
        r[o] = rr[q];
        u[o] = uu[q];
        m[o] = uu[q];
        z[o] = 0.0;
        qv[o] = 0.0;
        sv[o] = 0.0;
        pv[o] = 0.0;
    }
}

// The vector update of an iteration (INIT: of the start) from n = A m, with A m formed here
// from the real-space field emu and the reciprocal field's chunk partials (efdd = scale *
// sum + self term, as k_el_kfield_reduce).  A block takes CG_WPB waters (63 rows, one per
// lane): its four wavefronts first add a quarter of the field chunks each, in chunk order,
// and the first wavefront adds the quarters in order and updates the rows; the new w of
// each water is shared through LDS for m = M^-1 w.  Block partials (r.u, w.u,
// |polr2 r / r4pie0|^2, dipole count, max |m_i|) per water block.  INIT sets w = A u and
// leaves mu, r, u as they are.  The blocks after the waters' (Ewald sum over k) carry the
// dipole structure factors (INIT: S(u) = S(m), S(q) = S(p) = 0).
#define CG_WPB 7
template <bool INIT>
__global__ __launch_bounds__(256) void k_cg_update(ElecDev P, int nbw, const double* __restrict__ Fpart,
                                                   const double* __restrict__ emu, const double* __restrict__ minv,
                                                   double* __restrict__ m, double* __restrict__ dip,
                                                   double* __restrict__ r, double* __restrict__ u,
                                                   double* __restrict__ wv, double* __restrict__ z,
                                                   double* __restrict__ qv, double* __restrict__ sv,
                                                   double* __restrict__ pv, double* __restrict__ tmu,
                                                   const double* __restrict__ S, double* __restrict__ Sx,
                                                   double* __restrict__ Sp, double* __restrict__ Su,
                                                   double* __restrict__ Sqr, double* __restrict__ blk,
                                                   const double* __restrict__ state)
{
    if (state[2] != 0.0) return;
    const double alpha = INIT ? 0.0 : state[4], beta = INIT ? 0.0 : state[5];
    const int tid = threadIdx.x;
    if ((int)blockIdx.x >= nbw) {
        const long e = (long)(blockIdx.x - nbw) * 256 + tid;
        if (e >= 6L * P.K) return;
        const double sm = S[8 * (e / 6) + 2 + e % 6];
        if (INIT) {
            Su[e] = sm;
            Sqr[e] = 0.0;
            Sp[e] = 0.0;
        } else {
            const double sq = sm + beta * Sqr[e];
            const double sp = Su[e] + beta * Sp[e];
            Sqr[e] = sq;
            Sp[e] = sp;
            Sx[e] += alpha * sp;
            Su[e] -= alpha * sq;
        }
        return;
    }
    __shared__ double part[4][64];
    __shared__ double ws[64];
    __shared__ double wmax[4];
    const int lane = tid & 63, wave = tid >> 6;
    const int wl = lane / 9, i = lane - 9 * wl;
    const int w = blockIdx.x * CG_WPB + wl;
    const bool act = lane < 9 * CG_WPB && w < P.nw;
    const int o = 9 * w + i, j = 3 * w + i / 3, c = i % 3;
    const double p2 = act ? site_polr2(P, j) : 0.0;
    const bool on = act && p2 > P.toler;
    {
        // this wavefront's quarter of the field chunks
        double sum = 0.0;
        if (on && p2 > 1.e-6) {
            const int c0 = (P.fchunks * wave) / 4, c1 = (P.fchunks * (wave + 1)) / 4;
            const double* f = Fpart + (long)c * P.n + j;
            const long cs = 6L * P.n;
#pragma unroll 4
            for (int ch = c0; ch < c1; ++ch) sum += f[ch * cs];
        }
        part[wave][lane] = sum;
    }
    __syncthreads();
    double rr = 0.0, uu = 0.0, wn = 0.0;
    if (wave == 0 && on) {
        const double sum = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
        const double mm = m[o];
        const double nv = R4PIE0 * mm / p2 - (emu[o] + (p2 > 1.e-6 ? P.scale * sum : 0.0) + P.self_dd * mm);
        if (INIT) {
            wn = nv;
            wv[o] = wn;
            rr = r[o];
            uu = u[o];
        } else {
            const double zz = nv + beta * z[o];
            const double qq = mm + beta * qv[o];
            const double ss = wv[o] + beta * sv[o];
            const double pp = u[o] + beta * pv[o];
            z[o] = zz;
            qv[o] = qq;
            sv[o] = ss;
            pv[o] = pp;
            dip[o] += alpha * pp;
            tmu[o] += alpha * (R4PIE0 * pp / p2 - ss);
            rr = r[o] - alpha * ss;
            uu = u[o] - alpha * qq;
            wn = wv[o] - alpha * zz;
            r[o] = rr;
            u[o] = uu;
            wv[o] = wn;
        }
    }
    if (wave == 0) ws[lane] = wn;
    __syncthreads();
    double mn = 0.0;
    if (wave == 0 && on) {
        const double* mi = minv + 45L * w;
#pragma unroll
        for (int k = 0; k < 9; ++k) mn += mi[tri9(i, k)] * ws[9 * wl + k];
        m[o] = mn;
    }
    const double zs = p2 * rr / R4PIE0;
    const bool row = wave == 0 && on;
    double v[4] = {rr * uu, wn * uu, row ? zs * zs : 0.0, (row && c == 0) ? 1.0 : 0.0};
    block_sum_store(v, blk + 5 * blockIdx.x);
    // max |m_i| of the block (a maximum: the same in any order)
    double a = fabs(mn);
    for (int o2 = 32; o2 > 0; o2 >>= 1) a = fmax(a, __shfl_xor(a, o2, 64));
    if (lane == 0) wmax[wave] = a;
    __syncthreads();
    if (tid == 0) blk[5 * blockIdx.x + 4] = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
}

// The stopping test and scalars of iteration `it` alone (cg_step in one wavefront), for
// paths whose first kernel of the iteration does not take them.
__global__ __launch_bounds__(64) void k_cg_check(ElecDev P, int it, int nb, const double* __restrict__ blk,
                                                 double* __restrict__ state)
{
    if (state[2] != 0.0) return;
    (void)cg_step(P, it, nb, blk, state, true);
}


// ================================================================ final passes ======

// Polarization energy per site (coulomb_polar.f:58-103).
__global__ __launch_bounds__(256) void k_el_polar(ElecDev P, const double* __restrict__ efk,
                                                  const double* __restrict__ emu, const double* __restrict__ efdd,
                                                  const double* __restrict__ efdc, const double* __restrict__ dip,
                                                  double* __restrict__ ep, double* __restrict__ blk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (i < P.n) {
        double e = 0.0;
        const double p2 = site_polr2(P, i);
        if (p2 > 1.e-6) {
            double s = 0.0;
            for (int q = 0; q < 3; ++q)
                s += dip[3 * i + q] * (efk[3 * i + q] + efdc[3 * i + q] + 0.5 * (emu[3 * i + q] + efdd[3 * i + q]) + 0.0);
            const double edum = dip[3 * i] * dip[3 * i] + dip[3 * i + 1] * dip[3 * i + 1] + dip[3 * i + 2] * dip[3 * i + 2];
            e = -s;
            e = e + R4PIE0 * edum / p2 / 2.0;
        }
        ep[i] = e;
        v[0] = e;
    }
    block_sum_store(v, blk + blockIdx.x);
}

// Per-k scalars of the final reciprocal pass (ewald1p.f:696-730): akk (Qc^2 + Qs^2),
// akk |summu|^2 and the six omg components; block partials [nb][8].  Also stores, per k,
// the constants the site pass needs in kf[k][16].
__global__ __launch_bounds__(256) void k_el_kscal(ElecDev P, const double* __restrict__ kv,
                                                  const double* __restrict__ S, const double* __restrict__ Sq,
                                                  double* __restrict__ kf, double* __restrict__ blk)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (k < P.K) {
        const double kx = kv[8 * k], ky = kv[8 * k + 1], kz = kv[8 * k + 2], akk = kv[8 * k + 3];
        const double akv = kv[8 * k + 4];
        const double Qc = Sq[2 * k], Qs = Sq[2 * k + 1];
        const double Mc[3] = {S[8 * k + 2], S[8 * k + 3], S[8 * k + 4]};
        const double Ms[3] = {S[8 * k + 5], S[8 * k + 6], S[8 * k + 7]};
        const double kk[3] = {kx, ky, kz};
        const double skdmuckrs = kx * Mc[0] + ky * Mc[1] + kz * Mc[2];
        const double skdmuskrs = kx * Ms[0] + ky * Ms[1] + kz * Ms[2];
        const double summu1 = Qc - skdmuskrs, summu2 = Qs + skdmuckrs;
        v[0] = akk * (Qc * Qc + Qs * Qs);
        v[1] = akk * (summu1 * summu1 + summu2 * summu2);
        const double virprs = akv * (summu1 * summu1 + summu2 * summu2);
        const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
        for (int t = 0; t < 6; ++t)
            v[2 + t] = -virprs * kk[ia[t]] * kk[ib[t]] - 2.0 * akk * kk[ia[t]] * Ms[ib[t]] * summu1 +
                       2.0 * akk * kk[ia[t]] * Mc[ib[t]] * summu2;
        double* f = kf + 16 * k;
        f[0] = Qc;
        f[1] = Qs;
        f[2] = summu1;
        f[3] = summu2;
        f[4] = skdmuckrs;
        f[5] = skdmuskrs;
        f[6] = skdmuckrs;   // dpkMc
        f[7] = skdmuskrs;   // dpkMs
        for (int q = 0; q < 3; ++q) {
            f[8 + q] = Mc[q];
            f[11 + q] = Ms[q];
        }
    }
    block_sum_store(v, blk + 8 * blockIdx.x);
}


// Final reciprocal pass per site over a chunk of runs (ewald1p.f:550-693, 771-825):
// forces, ESP parts vesp_k / vesp_dc_k, per-atom energy and per-atom stress (scaled by
// akk, without the final prefactors).  The chunk's runs and per-k constants (k, akk,
// akv2, Qc, Qs, summu, k.Mc, k.Ms, Mc, Ms) are staged in LDS.  Partials
// Fpart[chunk][16][n]: f(3), vk, vdck, e, s(9).
#define EL_KF 17
__global__ __launch_bounds__(256) void k_el_kfinal(ElecDev P, const int* __restrict__ runs,
                                                   const int* __restrict__ runk, const double* __restrict__ phc,
                                                   const double* __restrict__ phs, const double* __restrict__ kv,
                                                   const double* __restrict__ kf, const double* __restrict__ chge,
                                                   const double* __restrict__ dip, double* __restrict__ Fpart)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x;
    const int j = blockIdx.x * 256 + tid;
    const int chunk = blockIdx.y;
    const int r0 = chunk * P.rpc, r1 = min(P.nrun, r0 + P.rpc);
    const int nr = max(0, r1 - r0);
    double* sk = lds;
    int* sr = reinterpret_cast<int*>(lds + (long)P.kmaxch * EL_KF);
    const int kbase = nr ? runk[r0] : 0;
    const int nk = nr ? ((r1 < P.nrun ? runk[r1] : P.K) - kbase) : 0;
    for (int idx = tid; idx < nk * EL_KF; idx += 256) {
        const int kk = idx / EL_KF, q = idx - kk * EL_KF;
        const long k = kbase + kk;
        double v;
        if (q < 4) v = kv[8 * k + q];                 // kx ky kz akk
        else if (q == 4) v = kv[8 * k + 5];           // akv2
        else if (q < 11) v = kf[16 * k + (q - 5)];    // Qc Qs summu1 summu2 k.Mc k.Ms
        else v = kf[16 * k + 8 + (q - 11)];           // Mc(3) Ms(3)
        sk[idx] = v;
    }
    for (int idx = tid; idx < nr * 4; idx += 256) sr[idx] = runs[4 * r0 + idx];
    __syncthreads();
    if (j >= P.n) return;
    const int ks = P.kstride;
    const bool pol = site_polr2(P, j) > 1.e-6;
    const double q = chge[j];
    const double mu[3] = {pol ? dip[3 * j] : 0.0, pol ? dip[3 * j + 1] : 0.0, pol ? dip[3 * j + 2] : 0.0};
    // Per k, with km = k.mu, the stress term of ewald1p.f:771-825,
    //   akk [B_ab (cc + cd + dd) + k_a X_b],  B_ab = delta_ab - akv2 k_a k_b,
    //   X_b = (s q + c km) Mc_b - (c q - s km) Ms_b + (c (Qs + k.Mc) + s (k.Ms - Qc)) mu_b,
    // is accumulated as delta_ab D - G_ab + Pm_ab + mu_b R_a with D = sum akk T, G_ab =
    // sum akk akv2 T k_a k_b, Pm_ab = sum k_a akk (..) and R_a = sum akk (..) k_a
    // (T = cc + cd + dd); the site energy sum akk q (c Qc + s Qs) is q times acc3.
    double f3[3] = {0.0, 0.0, 0.0}, vk_ = 0.0, vd_ = 0.0, D = 0.0;
    double G[6] = {0, 0, 0, 0, 0, 0}, Pm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, R[3] = {0, 0, 0};
    const double c1 = phc[((long)2 * ks + 1) * P.n + j], s1 = phs[((long)2 * ks + 1) * P.n + j];
    int kk = 0;
    for (int r = 0; r < nr; ++r) {
        const int l = sr[4 * r], m = sr[4 * r + 1], n0 = sr[4 * r + 2], cnt = sr[4 * r + 3];
        const int am = abs(m), an0 = abs(n0);
        double ckr, skr;
        run_phase(phc[(long)l * P.n + j], phs[(long)l * P.n + j], phc[((long)ks + am) * P.n + j],
                  phs[((long)ks + am) * P.n + j], phc[((long)2 * ks + an0) * P.n + j],
                  phs[((long)2 * ks + an0) * P.n + j], m, n0, ckr, skr);
        for (int t = 0; t < cnt; ++t) {
            const double* f = sk + EL_KF * (kk + t);
            const double kx = f[0], ky = f[1], kz = f[2];
            const double akk = f[3], akv2 = f[4];
            const double Qc = f[5], Qs = f[6], summu1 = f[7], summu2 = f[8];
            const double dpkMc = f[9], dpkMs = f[10];
            const double km = mu[0] * kx + mu[1] * ky + mu[2] * kz;
            const double ssmu = summu1 * (-skr * q - ckr * km) + summu2 * (-skr * km + ckr * q);
            const double wf = akk * ssmu;
            f3[0] -= kx * wf;
            f3[1] -= ky * wf;
            f3[2] -= kz * wf;
            const double cq = ckr * Qc + skr * Qs;
            vk_ += akk * cq;
            vd_ += akk * (skr * dpkMc - ckr * dpkMs);
            const double T = q * cq + q * (skr * dpkMc - ckr * dpkMs) + km * (ckr * Qs - skr * Qc) +
                             km * (ckr * dpkMc + skr * dpkMs);
            const double aT = akk * T;
            D += aT;
            const double u = aT * akv2;
            const double ux = u * kx, uy = u * ky, uz = u * kz;
            G[0] += ux * kx;
            G[1] += ux * ky;
            G[2] += ux * kz;
            G[3] += uy * ky;
            G[4] += uy * kz;
            G[5] += uz * kz;
            const double b1 = akk * (skr * q + ckr * km), b2 = akk * (ckr * q - skr * km);
            const double Y0 = b1 * f[11] - b2 * f[14], Y1 = b1 * f[12] - b2 * f[15], Y2 = b1 * f[13] - b2 * f[16];
            Pm[0] += kx * Y0;
            Pm[1] += kx * Y1;
            Pm[2] += kx * Y2;
            Pm[3] += ky * Y0;
            Pm[4] += ky * Y1;
            Pm[5] += ky * Y2;
            Pm[6] += kz * Y0;
            Pm[7] += kz * Y1;
            Pm[8] += kz * Y2;
            const double a3 = akk * (ckr * (Qs + dpkMc) + skr * (dpkMs - Qc));
            R[0] += a3 * kx;
            R[1] += a3 * ky;
            R[2] += a3 * kz;
            const double cn = ckr * c1 - skr * s1;
            skr = skr * c1 + ckr * s1;
            ckr = cn;
        }
        kk += cnt;
    }
    double acc[15];
    acc[0] = f3[0];
    acc[1] = f3[1];
    acc[2] = f3[2];
    acc[3] = vk_;
    acc[4] = vd_;
    acc[5] = q * vk_;
    const int gi[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            acc[6 + 3 * a + b] = (a == b ? D : 0.0) - G[gi[a][b]] + Pm[3 * a + b] + mu[b] * R[a];
    double* out = Fpart + (long)chunk * 16 * P.n;
#pragma unroll
    for (int t = 0; t < 15; ++t) out[(long)t * P.n + j] = acc[t];
}

// Sums of the final pass's site partials over the k chunks: ksum[t][j] for the 15 site
// quantities t of k_el_kfinal, 64 elements (t, j) per block.
__global__ __launch_bounds__(256) void k_el_ksum(ElecDev P, const double* __restrict__ Fpart,
                                                 double* __restrict__ ksum)
{
    const long e = (long)blockIdx.x * 64 + (threadIdx.x & 63);
    const bool valid = e < 15L * P.n;
    const double v = chunk_sum64(Fpart, 16L * P.n, P.kchunks, e, valid);
    if (threadIdx.x < 64 && valid) ksum[e] = v;
}

// Per-site results of the final reciprocal pass from the chunk sums, with the prefactors
// (ewald1p.f:941-1001): fk = scale * f, vesp_k, vesp_dc_k = scale * ..., e_ew1p = scal1 * e +
// self energy, s_ew1p = -scal1 * s, vesp_s.  Block partials [nb][2]: sum of q^2.
__global__ __launch_bounds__(256) void k_el_kfinal_reduce(ElecDev P, const double* __restrict__ ksum,
                                                          const double* __restrict__ chge, double* __restrict__ fk,
                                                          double* __restrict__ vk, double* __restrict__ vdck,
                                                          double* __restrict__ vs, double* __restrict__ e1,
                                                          double* __restrict__ s1, double* __restrict__ blk)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (j < P.n) {
        double a[15];
        for (int t = 0; t < 15; ++t) a[t] = ksum[(long)t * P.n + j];
        const double q = chge[j];
        for (int t = 0; t < 3; ++t) fk[3 * j + t] = P.scale * a[t];
        vk[j] = a[3] * P.scale;
        vdck[j] = a[4] * P.scale;
        vs[j] = -2.0 * R4PIE0 * q * P.alpha / SQRPI;
        e1[j] = -R4PIE0 * P.alpha * q * q / SQRPI + P.scal1 * a[5];
        for (int t = 0; t < 9; ++t) s1[9 * j + t] = -(P.scal1 * a[6 + t]);
        v[0] = q * q;
    }
    block_sum_store(v, blk + 2 * blockIdx.x);
}

// Real-space pair terms of one row (ewald2p.f:118-600; nthole = 4, lacs, lads): forces,
// per-atom energy, ESP parts, and the row's force-matrix contraction W_i (real partners
// only; for an M-site row the partner sums sum_j F and sum_j F (x) x_j, from which the
// M-row redistribution builds W of O, H1, H2).  Each pair is seen from both of its rows,
// so block partials carry half of each pair's energy, virial and stress.  One wavefront
// per site; partials [n][8]: energy, virial, stress (xx xy xz yy yz zz).
__global__ __launch_bounds__(64) void k_el_pairs(ElecDev P, const double* __restrict__ x,
                                                 const double* __restrict__ y, const double* __restrict__ z,
                                                 const double* __restrict__ chge, const double* __restrict__ dip,
                                                 const double* __restrict__ tab, const int* __restrict__ nbrA,
                                                 const int* __restrict__ cntA, double g34, double gln34,
                                                 double* __restrict__ f2, double* __restrict__ e2,
                                                 double* __restrict__ vr, double* __restrict__ vdcr,
                                                 double* __restrict__ w2, double* __restrict__ fmsum,
                                                 double* __restrict__ fmx, double* __restrict__ blk)
{
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    if (i >= P.n) return;
    const double qi = chge[i];
    const double pi_ = site_polr(P, i);
    const bool poli = site_polr2(P, i) > 1.e-6;
    const double mi[3] = {dip[3 * i], dip[3 * i + 1], dip[3 * i + 2]};
    const bool mrow = i >= P.nreal;
    double fsum[3] = {0, 0, 0}, esum = 0.0, vrs = 0.0, vdcs = 0.0, W[9], red[8];
#pragma unroll
    for (int t = 0; t < 9; ++t) W[t] = 0.0;
#pragma unroll
    for (int t = 0; t < 8; ++t) red[t] = 0.0;
    double fm[3] = {0, 0, 0};
    const int na = cntA[i];
    for (int s = lane; s < na; s += 64) {
        const int j = nbrA[(long)i * P.capA + s];
        double d[3];
        const double rsq = sep(P, x, y, z, i, j, d);
        if (!(P.rcut * P.rcut > rsq)) continue;
        const double rrr = sqrt(rsq);
        double bbb[4];
        table_interp(tab, P.drewd, rrr, bbb);
        const double rri = 1.0 / rrr, rsqi = 1.0 / rsq;
        const double r3i = 1.0 / (rrr * rrr * rrr);
        const double r5i = r3i * rsqi;
        const double qj = chge[j];
        const double cicj = qi * qj;
        const double edum = cicj * (bbb[0] + 1.0 / rrr) * R4PIE0;
        double espdum = (bbb[0] + 1.0 / rrr) * R4PIE0;
        double ffac = cicj * (r3i + bbb[1]) * R4PIE0;
        const double bb = pow(pi_ * site_polr(P, j), 1.0 / 6.0);
        const double u = rrr / bb, u4 = (u * u) * (u * u);
        double eng;
        {
            const double arob4 = P.ascc * u4;
            const double dum1 = exp(-arob4);
            const double sm = (dum1 - pow(arob4, 0.25) * g34 * gammq_nr(0.75, arob4, gln34)) / rrr * R4PIE0;
            eng = edum - cicj * sm;
            esum += 0.5 * edum;
            esum += -cicj * 0.5 * sm;
            espdum -= sm;
            ffac -= cicj * r3i * dum1 * R4PIE0;
        }
        const double mj[3] = {dip[3 * j], dip[3 * j + 1], dip[3 * j + 2]};
        const double rdmui = mi[0] * d[0] + mi[1] * d[1] + mi[2] * d[2];
        const double rdmuj = mj[0] * d[0] + mj[1] * d[1] + mj[2] * d[2];
        const double facmu0 = (qj * rdmui - qi * rdmuj) * (-3.0) * r5i * R4PIE0;
        const double facmui0 = r3i * qj * R4PIE0;
        const double facmuj0 = -r3i * qi * R4PIE0;
        double facmu = facmu0 - (qj * rdmui - qi * rdmuj) * bbb[2] * R4PIE0;
        double dfacmui = facmui0 + bbb[1] * qj * R4PIE0;
        double dfacmuj = facmuj0 - bbb[1] * qi * R4PIE0;
        double espdcdum = (bbb[1] + r3i) * R4PIE0;
        {
            const double arob4 = P.ascd * u4;
            const double dum1 = exp(-arob4);
            const double dum2 = (1.0 + (4.0 / 3.0) * arob4) * dum1;
            facmu -= facmu0 * dum2;
            dfacmui -= facmui0 * dum1;
            dfacmuj -= facmuj0 * dum1;
            espdcdum -= r3i * dum1 * R4PIE0;
        }
        if (poli && site_polr2(P, j) > 1.e-6) {
            const double dpidpj = mi[0] * mj[0] + mi[1] * mj[1] + mi[2] * mj[2];
            facmu += ((dpidpj * 3.0 + rdmui * rdmuj * rsqi * (-15.0)) * r5i + bbb[2] * dpidpj - bbb[3] * rdmui * rdmuj) *
                     R4PIE0;
            dfacmui += (3.0 * r5i + bbb[2]) * rdmuj * R4PIE0;
            dfacmuj += (3.0 * r5i + bbb[2]) * rdmui * R4PIE0;
            const double facmu1 = -15.0 * rdmui * rdmuj * rsqi * r5i * R4PIE0;
            const double facmu2 = 3.0 * dpidpj * r5i * R4PIE0;
            const double dfacmui1 = 3.0 * rdmuj * r5i * R4PIE0;
            const double dfacmuj1 = 3.0 * rdmui * r5i * R4PIE0;
            const double arob4 = P.athole * u4;
            const double dum1 = exp(-arob4);
            const double dum2 = (1.0 + (4.0 / 3.0) * arob4) * dum1;
            const double bb4 = (bb * bb) * (bb * bb);
            const double dum3 = 4.0 * P.athole * (4.0 * arob4 - 1.0) * dum1 * rsqi / bb4;
            const double dum4 = 4.0 * P.athole * dum1 * rrr / bb4;
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
        for (int q = 0; q < 3; ++q) fsum[q] += fv[q];
        vrs += qj * espdum;
        vdcs += rdmuj * espdcdum;
        red[0] += 0.5 * eng;
        red[1] += 0.5 * (-ffac * rsq - virdum);
        red[2] += 0.5 * d[0] * fv[0];
        red[3] += 0.5 * d[0] * fv[1];
        red[4] += 0.5 * d[0] * fv[2];
        red[5] += 0.5 * d[1] * fv[1];
        red[6] += 0.5 * d[1] * fv[2];
        red[7] += 0.5 * d[2] * fv[2];
        if (j < P.nreal) {
            if (!mrow) {
                const double dr[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) W[3 * a + b] += fv[a] * dr[b];
            } else {
                const double xj[3] = {x[j], y[j], z[j]};
                for (int a = 0; a < 3; ++a) {
                    fm[a] += fv[a];
                    for (int b = 0; b < 3; ++b) W[3 * a + b] += fv[a] * xj[b];
                }
            }
        }
    }
    for (int q = 0; q < 3; ++q) fsum[q] = wsum(fsum[q]);
    for (int q = 0; q < 3; ++q) fm[q] = wsum(fm[q]);
    for (int t = 0; t < 9; ++t) W[t] = wsum(W[t]);
    esum = wsum(esum);
    vrs = wsum(vrs);
    vdcs = wsum(vdcs);
    for (int t = 0; t < 8; ++t) red[t] = wsum(red[t]);
    if (lane == 0) {
        for (int q = 0; q < 3; ++q) f2[3 * i + q] = fsum[q];
        e2[i] = esum;
        vr[i] = vrs;
        vdcr[i] = vdcs;
        if (!mrow) {
            for (int t = 0; t < 9; ++t) w2[9 * i + t] = W[t];
        } else {
            for (int q = 0; q < 3; ++q) fmsum[3 * i + q] = fm[q];
            for (int t = 0; t < 9; ++t) fmx[9 * i + t] = W[t];
        }
        for (int t = 0; t < 8; ++t) blk[8 * i + t] = red[t];
    }
}

namespace {

// force-matrix entry F at (row a, column b) of water-local work: contraction into the
// per-atom tensors of the water's real atoms (M rows moved to O, H1, H2), columns of
// real atoms only, diagonal entries skipped (heatcurrent.f90:541-555, 711-719).
__device__ void fm_local(const ElecDev& P, const double* x, const double* y, const double* z, int a, int b,
                         const double* F, double* w3)
{
    if (b >= P.nreal) return;
    if (a < P.nreal) {
        if (a == b) return;
        const double dr[3] = {x[a] - x[b], y[a] - y[b], z[a] - z[b]};
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) w3[9 * a + 3 * p + q] += F[p] * dr[q];
        return;
    }
    const int w = a - P.nreal;
    const double wt[3] = {1.0 - P.gamma, 0.5 * P.gamma, 0.5 * P.gamma};
    for (int t = 0; t < 3; ++t) {
        const int r = 3 * w + t;
        if (r == b) continue;
        const double dr[3] = {x[r] - x[b], y[r] - y[b], z[r] - z[b]};
        for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) w3[9 * r + 3 * p + q] += wt[t] * F[p] * dr[q];
    }
}

}  // namespace

// Excluded-pair corrections of each water (ewald3p.f:80-391).  Writes the water's rows of
// f3, e3, vc, w3 (all zero-initialised by this kernel) and block partials [nb][8]:
// energy, virial, stress.
__global__ __launch_bounds__(64) void k_el_intra_final(ElecDev P, const double* __restrict__ x,
                                                       const double* __restrict__ y, const double* __restrict__ z,
                                                       const double* __restrict__ chge, const double* __restrict__ dip,
                                                       double* __restrict__ f3, double* __restrict__ e3,
                                                       double* __restrict__ vc, double* __restrict__ w3,
                                                       double* __restrict__ blk)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    double red[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (w < P.nw) {
        // the water's four sites (O, H1, H2, M); every sum of the water's rows is kept in
        // registers (slots 0..3, fully unrolled) and stored once
        const int s[4] = {3 * w, 3 * w + 1, 3 * w + 2, P.nreal + w};
        double xs[4][3];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            xs[a][0] = x[s[a]];
            xs[a][1] = y[s[a]];
            xs[a][2] = z[s[a]];
        }
        double fl[4][3], el[4], vl[4], wl[3][9];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            fl[a][0] = fl[a][1] = fl[a][2] = 0.0;
            el[a] = vl[a] = 0.0;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int t = 0; t < 9; ++t) wl[a][t] = 0.0;
        const double wt[3] = {1.0 - P.gamma, 0.5 * P.gamma, 0.5 * P.gamma};
        // force-matrix contraction of the pair force F on slot ai from slot bi (fm_local)
        auto fml = [&](int ai, int bi, const double (&F)[3]) {
            if (bi == 3) return;
            if (ai < 3) {
                if (ai == bi) return;
#pragma unroll
                for (int p = 0; p < 3; ++p)
#pragma unroll
                    for (int q = 0; q < 3; ++q) wl[ai][3 * p + q] += F[p] * (xs[ai][q] - xs[bi][q]);
                return;
            }
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                if (t == bi) continue;
#pragma unroll
                for (int p = 0; p < 3; ++p)
#pragma unroll
                    for (int q = 0; q < 3; ++q) wl[t][3 * p + q] += wt[t] * F[p] * (xs[t][q] - xs[bi][q]);
            }
        };
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = a + 1; b < 4; ++b) {
                const int i = s[a], j = s[b];
                double d[3];
                const double rsq = sep(P, x, y, z, i, j, d);
                const double rrr = sqrt(rsq);
                const double rri = 1.0 / rrr, rsqi = 1.0 / rsq;
                const double r3i = 1.0 / (rrr * rrr * rrr);
                const double r5i = r3i * rsqi;
                double bbb[4];
                bfun_direct(P.alpha, rrr, rsqi, bbb);
                const double qi = chge[i], qj = chge[j];
                const double cicj = qi * qj;
                red[0] += cicj * bbb[0] * R4PIE0;
                el[a] += 0.5 * cicj * bbb[0] * R4PIE0;
                el[b] += 0.5 * cicj * bbb[0] * R4PIE0;
                const double espdum = bbb[0] * R4PIE0;
                const double espdcdum = bbb[1] * R4PIE0;
                double ffac = cicj * bbb[1] * R4PIE0;
                const double mi[3] = {dip[3 * i], dip[3 * i + 1], dip[3 * i + 2]};
                const double mj[3] = {dip[3 * j], dip[3 * j + 1], dip[3 * j + 2]};
                const double rdmui = mi[0] * d[0] + mi[1] * d[1] + mi[2] * d[2];
                const double rdmuj = mj[0] * d[0] + mj[1] * d[1] + mj[2] * d[2];
                double facmu = (qj * rdmui - qi * rdmuj) * (-bbb[2]) * R4PIE0;
                double dfacmui = bbb[1] * qj * R4PIE0;
                double dfacmuj = -bbb[1] * qi * R4PIE0;
                if (site_polr2(P, i) > 1.e-6 && site_polr2(P, j) > 1.e-6) {
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
                    const double bb = pow(site_polr(P, i) * site_polr(P, j), 1.0 / 6.0);
                    const double ath = (i % 3 == 0 || j % 3 == 0) ? P.athole12 : P.athole13;
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
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    fl[a][q] += fv[q];
                    fl[b][q] -= fv[q];
                }
                const double mf[3] = {-fv[0], -fv[1], -fv[2]};
                fml(a, b, fv);
                fml(b, a, mf);
                red[1] += -ffac * rsq - virdum;
                red[2] += d[0] * fv[0];
                red[3] += d[0] * fv[1];
                red[4] += d[0] * fv[2];
                red[5] += d[1] * fv[1];
                red[6] += d[1] * fv[2];
                red[7] += d[2] * fv[2];
                vl[a] += qj * espdum;
                vl[b] += qi * espdum;
                vl[a] += rdmuj * espdcdum;
                vl[b] -= rdmui * espdcdum;
            }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int i = s[a];
#pragma unroll
            for (int q = 0; q < 3; ++q) f3[3 * i + q] = fl[a][q];
            e3[i] = el[a];
            vc[i] = vl[a];
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int t = 0; t < 9; ++t) w3[9L * s[a] + t] = wl[a][t];
    }
    block_sum_store(red, blk + 8 * blockIdx.x);
}

// ESP at every site (forces.f:876-880).
__global__ void k_el_vesp(int n, const double* __restrict__ vk, const double* __restrict__ vr,
                          const double* __restrict__ vs, const double* __restrict__ vc, const double* __restrict__ vdck,
                          const double* __restrict__ vdcr, double* __restrict__ vesp)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) vesp[i] = vk[i] + vr[i] + vs[i] + vc[i] + vdck[i] + vdcr[i] + 0.0;
}

// Charge-derivative forces of each water (qdforce.f90:70-566): adds to the water's rows
// of f3 and w3; block partials [nb][8]: 0, virial, stress.
__global__ __launch_bounds__(64) void k_el_qd(ElecDev P, const double* __restrict__ x, const double* __restrict__ y,
                                              const double* __restrict__ z, const double* __restrict__ dg,
                                              const double* __restrict__ vesp, double* __restrict__ f3,
                                              double* __restrict__ w3, double* __restrict__ blk)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    // [0..7] energy (unused), virial, stress6; [8..13] the angular terms of the per-atom qd
    // stresses (qdforce.f90:504-531, the ones recorded without HEAT_STRESS), summed
    double red[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (w < P.nw) {
        const int io = 3 * w, ih1 = io + 1, ih2 = io + 2, im = P.nreal + w;
        const int site[4] = {io, ih1, ih2, im};
        const double* G = dg + 6 * w;   // dp1dr1, dp1dr2, dp2dr1, dp2dr2, dp1dcos, dp2dcos
        double r1v[3], r2v[3];
        sep(P, x, y, z, ih1, io, r1v);
        sep(P, x, y, z, ih2, io, r2v);
        const double r1 = sqrt(r1v[0] * r1v[0] + r1v[1] * r1v[1] + r1v[2] * r1v[2]);
        const double r2 = sqrt(r2v[0] * r2v[0] + r2v[1] * r2v[1] + r2v[2] * r2v[2]);
        const double rdot = r1v[0] * r2v[0] + r1v[1] * r2v[1] + r1v[2] * r2v[2];
        const double cost = rdot / (r1 * r2);
        const double r3 = sqrt(r1 * r1 + r2 * r2 - 2.0 * r1 * r2 * cost);
        const double sint = sqrt(1.0 - cost * cost);
        const double cott = cost / sint;
        const double r1sqr = r1 * r1, r2sqr = r2 * r2;
        double dr[3][4][3];
        for (int k = 0; k < 3; ++k)
            for (int s = 0; s < 4; ++s)
                for (int q = 0; q < 3; ++q) dr[k][s][q] = 0.0;
        for (int q = 0; q < 3; ++q) {
            dr[0][1][q] = r1v[q] / r1;
            dr[0][0][q] = -dr[0][1][q];
            dr[1][2][q] = r2v[q] / r2;
            dr[1][0][q] = -dr[1][2][q];
            dr[2][1][q] = cott * (r1v[q] / r1sqr - r2v[q] / rdot);
            dr[2][2][q] = cott * (r2v[q] / r2sqr - r1v[q] / rdot);
            dr[2][0][q] = -dr[2][1][q] - dr[2][2][q];
        }
        const double tmp = P.gamma / 2.0 / (1.0 - P.gamma);
        double dq[4][4][3];
        for (int s = 0; s < 4; ++s)
            for (int q = 0; q < 3; ++q) {
                const double dp1 = G[0] * dr[0][s][q] + G[1] * dr[1][s][q] - sint * G[4] * dr[2][s][q];
                const double dp2 = G[2] * dr[0][s][q] + G[3] * dr[1][s][q] - sint * G[5] * dr[2][s][q];
                dq[0][s][q] = 0.0;
                dq[1][s][q] = dp1 * (tmp + 1.0) + dp2 * tmp;
                dq[2][s][q] = dp2 * (tmp + 1.0) + dp1 * tmp;
                dq[3][s][q] = -(dq[1][s][q] + dq[2][s][q]);
            }
        for (int j = 0; j < 4; ++j) {
            const double vt = vesp[site[j]];
            const int col = (j < 3) ? site[j] : (P.compat_qd ? P.nw + w : im);
            for (int s = 0; s < 4; ++s) {
                const double F[3] = {-dq[j][s][0] * vt, -dq[j][s][1] * vt, -dq[j][s][2] * vt};
                for (int q = 0; q < 3; ++q) f3[3 * site[s] + q] += F[q];
                fm_local(P, x, y, z, site[s], col, F, w3);
            }
        }
        double dqdr12[4] = {0, 0, 0, 0}, dqdr13[4] = {0, 0, 0, 0}, dqdcos[4] = {0, 0, 0, 0};
        dqdr12[1] = G[0] + (G[0] + G[2]) * tmp;
        dqdr13[1] = G[1] + (G[1] + G[3]) * tmp;
        dqdcos[1] = G[4] + (G[4] + G[5]) * tmp;
        dqdr12[2] = G[2] + (G[0] + G[2]) * tmp;
        dqdr13[2] = G[3] + (G[1] + G[3]) * tmp;
        dqdcos[2] = G[5] + (G[4] + G[5]) * tmp;
        dqdr12[3] = -(G[0] + G[2]) / (1.0 - P.gamma);
        dqdr13[3] = -(G[1] + G[3]) / (1.0 - P.gamma);
        dqdcos[3] = -(G[4] + G[5]) / (1.0 - P.gamma);
        double rv[2][3], r3v[3];
        sep(P, x, y, z, io, ih1, rv[0]);
        sep(P, x, y, z, io, ih2, rv[1]);
        sep(P, x, y, z, ih1, ih2, r3v);
        const double rr1 = sqrt(rv[0][0] * rv[0][0] + rv[0][1] * rv[0][1] + rv[0][2] * rv[0][2]);
        const double rr2 = sqrt(rv[1][0] * rv[1][0] + rv[1][1] * rv[1][1] + rv[1][2] * rv[1][2]);
        const double rr3 = sqrt(r3v[0] * r3v[0] + r3v[1] * r3v[1] + r3v[2] * r3v[2]);
        const double dcosdr12 = (r1 - r2 * cost) / r1 / r2;
        const double dcosdr13 = (r2 - r1 * cost) / r1 / r2;
        const double dcosdr23 = -r3 / r1 / r2;
        const int sa[6] = {0, 0, 0, 1, 1, 2}, sb[6] = {0, 1, 2, 1, 2, 2};
        for (int i = 0; i < 4; ++i) {
            const double vt = vesp[site[i]];
            for (int pr = 0; pr < 2; ++pr) {
                const double dqdr = pr == 0 ? dqdr12[i] : dqdr13[i];
                const double* r = rv[pr];
                const double rjk = pr == 0 ? rr1 : rr2;
                for (int t = 0; t < 6; ++t) red[2 + t] -= vt * dqdr * r[sa[t]] * r[sb[t]] / rjk;
                red[1] += vt * rjk * dqdr;
            }
            double dc[6];
            for (int t = 0; t < 6; ++t)
                dc[t] = dcosdr12 * rv[0][sa[t]] * rv[0][sb[t]] / rr1 + dcosdr13 * rv[1][sa[t]] * rv[1][sb[t]] / rr2 +
                        dcosdr23 * r3v[sa[t]] * r3v[sb[t]] / rr3;
            for (int t = 0; t < 6; ++t) red[2 + t] -= vt * dqdcos[i] * dc[t];
            for (int t = 0; t < 6; ++t) red[8 + t] += vt * dqdcos[i] * dc[t];
            red[1] += vt * (dc[0] + dc[3] + dc[5]) * dqdcos[i];
        }
    }
    block_sum_store(red, blk + 16 * blockIdx.x);
}

// Site forces and the per-real-atom outputs of the heat-current decomposition, with the
// M-site redistribution (dlpoly_cmd.f:2583-2606, heatcurrent.f90:302-555).  One thread
// per water: fout (3 per site; M forces moved to O, H1, H2 when distribute != 0),
// eout (energy ew1p + ew2p + ew3p + polar), sout (ew1p stress), wout (force-matrix
// contraction) for O, H1, H2.
__global__ void k_el_combine(ElecDev P, int distribute, const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, const double* __restrict__ fk, const double* __restrict__ f2,
                             const double* __restrict__ f3, const double* __restrict__ e1, const double* __restrict__ e2,
                             const double* __restrict__ e3, const double* __restrict__ ep,
                             const double* __restrict__ s1, const double* __restrict__ w2,
                             const double* __restrict__ w3, const double* __restrict__ fmsum,
                             const double* __restrict__ fmx, double* __restrict__ fout, double* __restrict__ eout,
                             double* __restrict__ sout, double* __restrict__ wout)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= P.nw) return;
    const int m = P.nreal + w;
    const double wt[3] = {1.0 - P.gamma, P.gamma / 2.0, P.gamma / 2.0};
    double fm[3];
    for (int q = 0; q < 3; ++q) fm[q] = fk[3 * m + q] + f2[3 * m + q] + f3[3 * m + q];
    const double em = e1[m] + e2[m] + e3[m] + ep[m];
    for (int t = 0; t < 3; ++t) {
        const int r = 3 * w + t;
        for (int q = 0; q < 3; ++q) {
            const double f = fk[3 * r + q] + f2[3 * r + q] + f3[3 * r + q];
            fout[3 * r + q] = distribute ? f + wt[t] * fm[q] : f;
        }
        eout[r] = (e1[r] + e2[r] + e3[r] + ep[r]) + wt[t] * em;
        for (int q = 0; q < 9; ++q) sout[9 * r + q] = s1[9 * r + q] + wt[t] * s1[9 * m + q];
        const double xr[3] = {x[r], y[r], z[r]};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                wout[9 * r + 3 * a + b] =
                    w2[9 * r + 3 * a + b] + w3[9 * r + 3 * a + b] + wt[t] * (fmsum[3 * m + a] * xr[b] - fmx[9 * m + 3 * a + b]);
    }
    for (int q = 0; q < 3; ++q) fout[3 * m + q] = distribute ? 0.0 : fm[q];
}

// Fixed-order sums of block partials: set b has nb[b] blocks of stride[b] values, of which
// nv[b] are summed into out[off[b] ..].  One block of 256 threads per (set, value), listed
// in sv[]: thread t adds blocks t, t + 256, ..., then the wavefront sums and the four wave
// sums are added in order.
struct SumSets {
    const double* blk[7];
    int nb[7], stride[7], nv[7], off[7];
    int nsv;
    unsigned char sv[64][2];   // (set, value) of each block
};
__global__ __launch_bounds__(256) void k_el_sum_sets(SumSets S, double* __restrict__ out)
{
    __shared__ double ws[4];
    const int b = S.sv[blockIdx.x][0], q = S.sv[blockIdx.x][1];
    double s = 0.0;
    for (int k = threadIdx.x; k < S.nb[b]; k += 256) s += S.blk[b][(long)k * S.stride[b] + q];
    s = wsum(s);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[S.off[b] + q] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
}

// Zero a double array (graph-safe replacement for a memset node).
__global__ void k_el_zero(double* __restrict__ p, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ============================================================ host launchers ======

namespace {
inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }
// dynamic LDS of the reciprocal-space passes (bytes)
inline size_t sfac_lds(const ElecDev& P)
{
    // the tile of table rows and site weights, reused afterwards for the site-group sums
    const size_t tile = 2 * EL_TROW * (size_t)(P.km[0] + P.km[1] + P.km[2] + 3) + 4 * EL_TILE;
    const size_t red = (size_t)(EL_SG - 1) * 64 * 8;
    return sizeof(double) * (tile > red ? tile : red);
}
inline size_t kfield_lds(const ElecDev& P)
{
    // per-k constants (12), runs (4 ints) and run offsets (1 int), then the run-group sums
    const size_t stage = (size_t)P.fkmax * 12 + (5 * (size_t)P.frpc + 1) / 2 + 1;
    const size_t red = (size_t)(EL_KG - 1) * 64 * 12;
    return sizeof(double) * (stage > red ? stage : red);
}
inline size_t kfinal_lds(const ElecDev& P) { return sizeof(double) * ((size_t)P.kmaxch * EL_KF + 2 * (size_t)P.rpc + 2); }
}

int el_launch_setup(ElecBuffers& E, const ElecDev& P, CellList& cells, hipStream_t s)
{
    const int nwb = ceil_div(P.nw, 64);
    hipLaunchKernelGGL(k_el_sites, dim3(nwb), dim3(64), 0, s, P, E.x, E.y, E.z, E.chge, E.dg);
    if (P.cells.on) {
        const int rc = cells_build(cells, P.n, E.x, E.y, E.z, 1, s);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_el_nbr, dim3(P.n), dim3(64), 0, s, P, E.x, E.y, E.z, E.nbrA, E.cntA, E.nbrP, E.cntP, E.err);
    return (int)hipGetLastError();
}

// after the neighbour counts were checked by the host
// The reciprocal field of the dipoles mu at the first n sites (particle mesh) into
// Fpart[c * P.n + j] as a k-sum partial (times kap), for the chunk-sum readers with one chunk
// (scale_state > 0: mu spread times the power of two in state[scale_state], see SpSpec)
static void el_spme_dipole_field(ElecBuffers& E, const ElecSpme& sp, int n, const double* mu, const double* state,
                                 hipStream_t s, int scale_state = 0)
{
    SpSpec spec{};
    spec.ngr = 1;
    spec.mu_on[0] = 1;
    spec.comp[0] = -1;
    spec.scale_state = scale_state;
    double2* g = sp.B->grid;
    sp_launch_spread(*sp.B, *sp.G, n, nullptr, mu, spec, g, true, state, s);
    sp_launch_convolve(*sp.B, *sp.G, g, 1, true, state, s);
    sp_launch_field(*sp.B, *sp.G, n, g, 1, sp.kap, E.Fpart, E.n, state, s, scale_state);
}

static inline ElecDev one_chunk(const ElecDev& P)
{
    ElecDev Q = P;
    Q.fchunks = 1;
    Q.K = 0;
    return Q;
}

int el_launch_prep(ElecBuffers& E, const ElecDev& P, const double* rcell_dev, const ElecSpme* sp, hipStream_t s)
{
    const int nwb = ceil_div(P.nw, 64);
    hipLaunchKernelGGL(k_el_prep, dim3(P.n), dim3(64), 0, s, P, E.x, E.y, E.z, E.chge, E.tab, E.nbrA, E.cntA, E.nbrP,
                       E.cntP, E.tens, E.efk);
    hipLaunchKernelGGL(k_el_intra_prep, dim3(nwb), dim3(64), 0, s, P, E.x, E.y, E.z, E.chge, E.tin, E.efk);
    if (sp && !sp->B) {
        // no reciprocal space (an isolated cluster): zero fields, in the layout of the k-sum
        // field partials (one chunk), so that the reductions form efdd (the self term is 0
        // without an Ewald split) and efdc
        el_launch_zero(E.Fpart, 6L * P.n, s);
        hipLaunchKernelGGL(k_el_kfield_reduce<true>, dim3(ceil_div(6L * P.n, 64)), dim3(256), 0, s, one_chunk(P),
                           E.Fpart, E.dip, E.efdd, E.efdc, E.state);
        return (int)hipGetLastError();
    }
    if (sp) {
        // particle mesh: splines of every site, then the fields of the start dipoles (grid 0)
        // and of the charges (grid 1) at every site, in the layout of the k-sum field partials
        sp_launch_prep(*sp->B, *sp->G, P.n, E.x, E.y, E.z, s);
        SpSpec spec{};
        spec.ngr = 2;
        spec.mu_on[0] = 1;
        spec.comp[0] = -1;
        spec.q_on[1] = 1;
        spec.comp[1] = -1;
        double2* g = sp->B->grid;
        sp_launch_spread(*sp->B, *sp->G, P.n, E.chge, E.dip, spec, g, true, nullptr, s);
        sp_launch_convolve(*sp->B, *sp->G, g, 2, true, nullptr, s);
        sp_launch_field(*sp->B, *sp->G, P.n, g, 2, sp->kap, E.Fpart, P.n, nullptr, s);
        hipLaunchKernelGGL(k_el_kfield_reduce<true>, dim3(ceil_div(6L * P.n, 64)), dim3(256), 0, s, one_chunk(P),
                           E.Fpart, E.dip, E.efdd, E.efdc, E.state);
        return (int)hipGetLastError();
    }
    hipLaunchKernelGGL(k_el_phase, dim3(ceil_div(P.n, 64)), dim3(64), 0, s, P, E.x, E.y, E.z, rcell_dev, E.phc,
                       E.phs);
    // charge structure factors and the iloop = 1 reciprocal fields
    hipLaunchKernelGGL(k_el_sfac<true>, dim3(ceil_div((long)EL_SPLIT * P.nrun, 64), E.schunks), dim3(64 * EL_SG), sfac_lds(P), s, P, E.spc,
                       E.runs, E.runk, E.phc, E.phs, E.chge, E.dip, E.Spart, E.state, nullptr, 0, 0);
    hipLaunchKernelGGL(k_el_sreduce<true>, dim3(ceil_div(8L * P.K, 64)), dim3(256), 0, s, P, E.schunks, E.Spart, E.kv, E.S,
                       E.Sq, E.kc, E.state);
    hipLaunchKernelGGL(k_el_kfield<true>, dim3(ceil_div(P.n, 128), P.fchunks), dim3(64 * EL_KG), kfield_lds(P), s, P, E.runs,
                       E.runk, E.phc, E.phs, E.kc, E.Fpart, E.state);
    hipLaunchKernelGGL(k_el_kfield_reduce<true>, dim3(ceil_div(6L * P.n, 64)), dim3(256), 0, s, P, E.Fpart, E.dip, E.efdd,
                       E.efdc, E.state);
    return (int)hipGetLastError();
}

int el_launch_iteration(ElecBuffers& E, const ElecDev& P, int first, const ElecSpme* sp, hipStream_t s)
{
    if (!first && sp && !sp->B) {
        // no reciprocal space: efdd keeps the preparation's zeros
    } else if (!first && sp) {
        el_spme_dipole_field(E, *sp, P.nrow, E.dip, E.state, s);
        hipLaunchKernelGGL(k_el_kfield_reduce<false>, dim3(ceil_div(3L * P.nrow, 64)), dim3(256), 0, s, one_chunk(P),
                           E.Fpart, E.dip, E.efdd, E.efdc, E.state);
    } else if (!first) {
        // reciprocal dipole field from the current dipoles (the first iteration's came
        // with the charge field)
        hipLaunchKernelGGL(k_el_sfac<false>, dim3(ceil_div((long)EL_SPLIT * P.nrun, 64), E.dchunks), dim3(64 * EL_SG), sfac_lds(P), s, P,
                           E.dspc, E.runs, E.runk, E.phc, E.phs, E.chge, E.dip, E.Spart, E.state, nullptr, 0, 0);
        hipLaunchKernelGGL(k_el_sreduce<false>, dim3(ceil_div(6L * P.K, 60)), dim3(256), 0, s, P, E.dchunks, E.Spart, E.kv,
                           E.S, E.Sq, E.kc, E.state);
        hipLaunchKernelGGL(k_el_kfield<false>, dim3(ceil_div(P.nrow, 128), P.fchunks), dim3(64 * EL_KG), kfield_lds(P), s, P,
                           E.runs, E.runk, E.phc, E.phs, E.kc, E.Fpart, E.state);
        hipLaunchKernelGGL(k_el_kfield_reduce<false>, dim3(ceil_div(3L * P.nrow, 64)), dim3(256), 0, s, P, E.Fpart, E.dip,
                           E.efdd, E.efdc, E.state);
    }
    hipLaunchKernelGGL(k_el_emu, dim3(ceil_div(P.nrow, 4)), dim3(256), 0, s, P, E.nbrP, E.cntP, E.tens, E.tin, E.dip,
                       E.emu, E.state, nullptr, 0, 0);
    const int nb = ceil_div(P.nrow, 256);
    hipLaunchKernelGGL(k_el_sor, dim3(nb), dim3(256), 0, s, P, E.efk, E.emu, E.efdd, E.efdc, E.dip, E.blk, E.state);
    hipLaunchKernelGGL(k_el_converge, dim3(1), dim3(64), 0, s, P, nb, E.blk, E.state);
    return (int)hipGetLastError();
}

// The dipole-field operator applied to the solver's m (emu: real-space rows; Fpart: the
// reciprocal field's partials), for the k-sum: structure factors (taking the stopping test
// of iteration `it` first when `check`), their sum and emu in one launch, the field.
static void el_cg_apply_ksum(ElecBuffers& E, const ElecDev& P, bool check, int it, hipStream_t s)
{
    const int nbw = ceil_div(P.nw, CG_WPB), nsr = ceil_div(6 * P.K, 60);
    hipLaunchKernelGGL(k_el_sfac<false>, dim3(ceil_div((long)EL_SPLIT * P.nrun, 64), E.dchunks), dim3(64 * EL_SG),
                       sfac_lds(P), s, P, E.dspc, E.runs, E.runk, E.phc, E.phs, E.chge, E.cg_m, E.Spart, E.state,
                       check ? E.cgblk : nullptr, nbw, it);
    hipLaunchKernelGGL(k_el_sreduce_emu, dim3(nsr + ceil_div(P.nrow, 4)), dim3(256), 0, s, P, nsr, E.dchunks, E.Spart,
                       E.kv, E.S, E.Sq, E.kc, E.nbrP, E.cntP, E.tens, E.tin, E.cg_m, E.emu, E.state);
    hipLaunchKernelGGL(k_el_kfield<false>, dim3(ceil_div(P.nrow, 128), P.fchunks), dim3(64 * EL_KG), kfield_lds(P), s, P,
                       E.runs, E.runk, E.phc, E.phs, E.kc, E.Fpart, E.state);
}

// The same for the particle mesh (emu first, with the stopping test when `check`) and for no
// reciprocal space (emu alone; the preparation's zero field partials stand)
static void el_cg_apply_other(ElecBuffers& E, const ElecDev& P, const ElecSpme& sp, bool check, int it, hipStream_t s)
{
    const int nbw = ceil_div(P.nw, CG_WPB);
    hipLaunchKernelGGL(k_el_emu, dim3(ceil_div(P.nrow, 4)), dim3(256), 0, s, P, E.nbrP, E.cntP, E.tens, E.tin, E.cg_m,
                       E.emu, E.state, check ? E.cgblk : nullptr, nbw, it);
    if (sp.B) el_spme_dipole_field(E, sp, P.nrow, E.cg_m, E.state, s, 9);
}

static void el_cg_update(ElecBuffers& E, const ElecDev& P, const ElecSpme* sp, bool init, hipStream_t s)
{
    const int nbw = ceil_div(P.nw, CG_WPB);
    const ElecDev Q = sp ? one_chunk(P) : P;
    const int nsb = sp ? 0 : ceil_div(6 * P.K, 256);
    if (init)
        hipLaunchKernelGGL(k_cg_update<true>, dim3(nbw + nsb), dim3(256), 0, s, Q, nbw, E.Fpart, E.emu, E.cg_minv,
                           E.cg_m, E.dip, E.cg_r, E.cg_u, E.cg_w, E.cg_z, E.cg_q, E.cg_s, E.cg_p, E.cg_tmu, E.S,
                           E.cg_smu, E.cg_sp, E.cg_su, E.cg_sq, E.cgblk, E.state);
    else
        hipLaunchKernelGGL(k_cg_update<false>, dim3(nbw + nsb), dim3(256), 0, s, Q, nbw, E.Fpart, E.emu, E.cg_minv,
                           E.cg_m, E.dip, E.cg_r, E.cg_u, E.cg_w, E.cg_z, E.cg_q, E.cg_s, E.cg_p, E.cg_tmu, E.S,
                           E.cg_smu, E.cg_sp, E.cg_su, E.cg_sq, E.cgblk, E.state);
}

// conjugate-gradient start: residual of the start dipoles (their fields came with the
// charge fields of the preparation, plus the real-space dipole field here), u = M^-1 r, and
// the first application of A, to u, for w = A u
int el_launch_cg_init(ElecBuffers& E, const ElecDev& P, const ElecSpme* sp, hipStream_t s)
{
    const int nbw = ceil_div(P.nw, 256);
    hipLaunchKernelGGL(k_cg_block, dim3(nbw), dim3(256), 0, s, P, E.x, E.y, E.z, E.cg_minv);
    hipLaunchKernelGGL(k_el_emu, dim3(ceil_div(P.nrow, 4)), dim3(256), 0, s, P, E.nbrP, E.cntP, E.tens, E.tin, E.dip,
                       E.emu, E.state, nullptr, 0, 0);
    hipLaunchKernelGGL(k_cg_init, dim3(nbw), dim3(256), 0, s, P, E.efk, E.efdc, E.emu, E.efdd, E.dip, E.cg_minv, E.cg_r,
                       E.cg_u, E.cg_m, E.cg_z, E.cg_q, E.cg_s, E.cg_p, E.cg_tmu);
    if (!sp) {
        hipLaunchKernelGGL(k_cg_smu_init, dim3(ceil_div(6L * P.K, 256)), dim3(256), 0, s, P, E.S, E.cg_smu);
        el_cg_apply_ksum(E, P, false, 0, s);
    } else {
        el_cg_apply_other(E, P, *sp, false, 0, s);
    }
    el_cg_update(E, P, sp, true, s);
    return (int)hipGetLastError();
}

// Iteration `it` (0-based): the stopping test and scalars from the last update's partials,
// n = A m (real-space rows, structure factors, reciprocal field), then the vector update.
// With the particle mesh or no reciprocal space the structure factors are not carried.
int el_launch_cg_iteration(ElecBuffers& E, const ElecDev& P, int it, const ElecSpme* sp, hipStream_t s)
{
    if (sp) el_cg_apply_other(E, P, *sp, true, it, s);
    else el_cg_apply_ksum(E, P, true, it, s);
    el_cg_update(E, P, sp, false, s);
    return (int)hipGetLastError();
}

// `reps` launches of the dipole structure-factor pass over the current dipoles, bracketed
// by events on the stream (timing only: the partials it writes are rewritten by every
// evaluation before use)
int el_time_sfac(ElecBuffers& E, const ElecDev& P, int reps, hipEvent_t a, hipEvent_t b, hipStream_t s)
{
    (void)hipEventRecord(a, s);
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_el_sfac<false>, dim3(ceil_div((long)EL_SPLIT * P.nrun, 64), E.dchunks), dim3(64 * EL_SG), sfac_lds(P), s, P,
                           E.dspc, E.runs, E.runk, E.phc, E.phs, E.chge, E.dip, E.Spart, E.zero_state, nullptr, 0, 0);
    (void)hipEventRecord(b, s);
    return (int)hipGetLastError();
}

int el_launch_final(ElecBuffers& E, const ElecDev& P, int distribute, double g34, double gln34, int refresh,
                    const ElecSpme* sp, hipStream_t s)
{
    // fields and structure factors of the final dipoles: carried along by the conjugate-
    // gradient solver; the SOR path keeps those of its last iteration, as the reference does
    // (the particle mesh carries no structure factors)
    if (refresh) {
        const ElecDev Q = sp ? one_chunk(P) : P;
        hipLaunchKernelGGL(k_cg_final_fields, dim3(ceil_div(std::max(3L * Q.nrow, 6L * Q.K), 256)), dim3(256), 0, s, Q,
                           E.cg_tmu, E.cg_smu, E.emu, E.efdd, E.S);
    }
    // k-sum: one partial per 256 k; particle mesh: one per 4 sites (k_sp_elec_final's blocks)
    const int nsb = ceil_div(P.n, 256), nwb = ceil_div(P.nw, 64), nkb = sp ? ceil_div(P.n, 4) : ceil_div(P.K, 256);
    double* b_pol = E.blk;                       // [nsb]
    double* b_ks = b_pol + nsb;                  // [nkb][8]
    double* b_q2 = b_ks + 8L * nkb;              // [nsb][2]
    double* b_p2 = b_q2 + 2L * nsb;              // [n][8]
    double* b_p3 = b_p2 + 8L * P.n;              // [nwb][8]
    double* b_qd = b_p3 + 8L * nwb;              // [nwb][16]
    hipLaunchKernelGGL(k_el_polar, dim3(nsb), dim3(256), 0, s, P, E.efk, E.emu, E.efdd, E.efdc, E.dip, E.ep, b_pol);
    if (sp && !sp->B) {
        // no reciprocal space: zero site sums and per-block partials for the k-sum's final
        // kernel (forces, ESP, per-atom energy and stress of the reciprocal part all 0)
        double* d[2] = {E.ksum, b_ks};
        const double* sr[2] = {nullptr, nullptr};
        const long len[2] = {15L * P.n, 8L * nkb};
        el_launch_copy_segs(2, d, sr, len, s);
    } else if (sp) {
        // particle mesh: grids 0..4 hold q + mu, q, and mu_x, mu_y, mu_z as charges; the
        // stress kernels of grid 0 go to grids 5..10; every per-site quantity of the k-sum
        // site pass then comes from one gather (k_sp_elec_final), in the k-sum pass's units
        SpSpec spec{};
        spec.ngr = 5;
        spec.q_on[0] = 1;
        spec.mu_on[0] = 1;
        spec.q_on[1] = 1;
        spec.comp[0] = spec.comp[1] = -1;
        for (int b = 0; b < 3; ++b) spec.comp[2 + b] = (signed char)b;
        SpmeBuffers& B = *sp->B;
        const SpmeGeom& G = *sp->G;
        double2* g = B.grid;
        sp_launch_spread(B, G, P.n, E.chge, E.dip, spec, g, true, nullptr, s);
        sp_launch_fft(B, G, g, 5, false, nullptr, s, true);
        sp_launch_influence_stress(B, G, g, g + 5 * B.ng_cap, s);
        sp_launch_influence(B, G, g + B.ng_cap, 4, false, nullptr, s);
        sp_launch_fft(B, G, g, 11, true, nullptr, s, false);
        sp_launch_elec_final(B, G, P.n, E.chge, E.dip, g, sp->kap, E.ksum, b_ks, s);
    } else if (!refresh) {
        // structure factors of the converged dipoles (already computed with the refreshed fields)
        hipLaunchKernelGGL(k_el_sfac<false>, dim3(ceil_div((long)EL_SPLIT * P.nrun, 64), E.dchunks), dim3(64 * EL_SG), sfac_lds(P), s, P,
                           E.dspc, E.runs, E.runk, E.phc, E.phs, E.chge, E.dip, E.Spart, E.zero_state, nullptr, 0, 0);
        hipLaunchKernelGGL(k_el_sreduce<false>, dim3(ceil_div(6L * P.K, 60)), dim3(256), 0, s, P, E.dchunks, E.Spart, E.kv, E.S, E.Sq, E.kc,
                           (const double*)E.zero_state);
    }
    if (!sp) {
        hipLaunchKernelGGL(k_el_kscal, dim3(nkb), dim3(256), 0, s, P, E.kv, E.S, E.Sq, E.kf, b_ks);
        hipLaunchKernelGGL(k_el_kfinal, dim3(ceil_div(P.n, 256), P.kchunks), dim3(256), kfinal_lds(P), s, P, E.runs,
                           E.runk, E.phc, E.phs, E.kv, E.kf, E.chge, E.dip, E.Fpart);
        hipLaunchKernelGGL(k_el_ksum, dim3(ceil_div(15L * P.n, 64)), dim3(256), 0, s, P, E.Fpart, E.ksum);
    }
    hipLaunchKernelGGL(k_el_kfinal_reduce, dim3(nsb), dim3(256), 0, s, P, E.ksum, E.chge, E.fk, E.vk, E.vdck, E.vs,
                       E.e1, E.s1, b_q2);
    hipLaunchKernelGGL(k_el_pairs, dim3(P.n), dim3(64), 0, s, P, E.x, E.y, E.z, E.chge, E.dip, E.tab, E.nbrA, E.cntA,
                       g34, gln34, E.f2, E.e2, E.vr, E.vdcr, E.w2, E.fmsum, E.fmx, b_p2);
    hipLaunchKernelGGL(k_el_intra_final, dim3(nwb), dim3(64), 0, s, P, E.x, E.y, E.z, E.chge, E.dip, E.f3, E.e3, E.vc,
                       E.w3, b_p3);
    hipLaunchKernelGGL(k_el_vesp, dim3(nsb), dim3(256), 0, s, P.n, E.vk, E.vr, E.vs, E.vc, E.vdck, E.vdcr, E.vesp);
    hipLaunchKernelGGL(k_el_qd, dim3(nwb), dim3(64), 0, s, P, E.x, E.y, E.z, E.dg, E.vesp, E.f3, E.w3, b_qd);
    hipLaunchKernelGGL(k_el_combine, dim3(nwb), dim3(64), 0, s, P, distribute, E.x, E.y, E.z, E.fk, E.f2, E.f3, E.e1,
                       E.e2, E.e3, E.ep, E.s1, E.w2, E.w3, E.fmsum, E.fmx, E.fsite, E.eout, E.sout, E.wout);
    // scalars: [0] eng_polar, [1..8] k sums (engc, engdum, omg6), [9..10] sum q^2,
    // [11..18] pairs (energy, virial, stress6), [19..26] excluded pairs, [27..34] qd,
    // [35..40] angular per-atom qd stresses
    SumSets ss{};
    const double* sb[7] = {b_pol, b_ks, b_q2, b_p2, b_p3, b_qd, b_qd + 8};
    const int snb[7] = {nsb, nkb, nsb, P.n, nwb, nwb, nwb}, sst[7] = {1, 8, 2, 8, 8, 16, 16};
    const int snv[7] = {1, 8, 2, 8, 8, 8, 6}, soff[7] = {0, 1, 9, 11, 19, 27, 35};
    ss.nsv = 0;
    for (int q = 0; q < 7; ++q) {
        ss.blk[q] = sb[q];
        ss.nb[q] = snb[q];
        ss.stride[q] = sst[q];
        ss.nv[q] = snv[q];
        ss.off[q] = soff[q];
        for (int v = 0; v < snv[q]; ++v) {
            ss.sv[ss.nsv][0] = (unsigned char)q;
            ss.sv[ss.nsv][1] = (unsigned char)v;
            ++ss.nsv;
        }
    }
    hipLaunchKernelGGL(k_el_sum_sets, dim3(ss.nsv), dim3(256), 0, s, ss, E.scal);
    return (int)hipGetLastError();
}

void el_launch_zero(double* p, long n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_el_zero, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n);
}

__global__ void k_el_add(double* __restrict__ dst, const double* __restrict__ src, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] += src[i];
}

// dst[i] += src[i], i < n
void el_launch_add(double* dst, const double* src, long n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_el_add, dim3(ceil_div(n, 256)), dim3(256), 0, s, dst, src, n);
}

__global__ void k_el_copy(double* __restrict__ dst, const double* __restrict__ src, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_el_pack(const double* __restrict__ state, const int* __restrict__ err,
                          const double* __restrict__ scal, const int* __restrict__ spflag, double* __restrict__ out)
{
    const int t = threadIdx.x;
    if (t >= EL_PK_N) return;
    double v;
    if (t < EL_PK_ERR) v = state[t - EL_PK_ST];
    else if (t < EL_PK_SCAL) v = (double)err[0];
    else if (t < EL_PK_SPF) v = scal[t - EL_PK_SCAL];
    else v = spflag ? (double)spflag[t - EL_PK_SPF] : 0.0;
    out[t] = v;
}

void el_launch_pack(const double* state, const int* err, const double* scal, const int* spflag, double* out,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_el_pack, dim3(1), dim3(64), 0, s, state, err, scal, spflag, out);
}

// Up to EL_NSEG copies dst_k[i] = src_k[i], i < n_k (a null source clears dst_k) in one
// launch, segment k in blockIdx.y: the small copies around an evaluation as one graph node.
struct ElCopySegs {
    double* dst[EL_NSEG];
    const double* src[EL_NSEG];
    long n[EL_NSEG];
};

__global__ void k_el_copy_segs(ElCopySegs S)
{
    const int k = blockIdx.y;
    double* __restrict__ d = S.dst[k];
    const double* __restrict__ src = S.src[k];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < S.n[k]; i += (long)gridDim.x * blockDim.x)
        d[i] = src ? src[i] : 0.0;
}

void el_launch_copy_segs(int nseg, double* const* dst, const double* const* src, const long* n, hipStream_t s)
{
    ElCopySegs S{};
    long nmax = 0;
    for (int k = 0; k < nseg; ++k) {
        S.dst[k] = dst[k];
        S.src[k] = src[k];
        S.n[k] = n[k];
        nmax = n[k] > nmax ? n[k] : nmax;
    }
    if (nseg <= 0 || nmax <= 0) return;
    const long blocks = ceil_div(nmax, 256) < 1024 ? ceil_div(nmax, 256) : 1024;
    hipLaunchKernelGGL(k_el_copy_segs, dim3(blocks, nseg), dim3(256), 0, s, S);
}

// Up to EL_NSEG sums dst_k[i] = a_k[i] + b_k[i], i < n_k, in one launch (segment k in
// blockIdx.y).
struct ElSumSegs {
    double* dst[EL_NSEG];
    const double* a[EL_NSEG];
    const double* b[EL_NSEG];
    long n[EL_NSEG];
};

__global__ void k_el_sum_segs(ElSumSegs S)
{
    const int k = blockIdx.y;
    double* __restrict__ d = S.dst[k];
    const double* __restrict__ a = S.a[k];
    const double* __restrict__ b = S.b[k];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < S.n[k]; i += (long)gridDim.x * blockDim.x)
        d[i] = a[i] + b[i];
}

void el_launch_sum_segs(int nseg, double* const* dst, const double* const* a, const double* const* b, const long* n,
                        hipStream_t s)
{
    ElSumSegs S{};
    long nmax = 0;
    for (int k = 0; k < nseg; ++k) {
        S.dst[k] = dst[k];
        S.a[k] = a[k];
        S.b[k] = b[k];
        S.n[k] = n[k];
        nmax = n[k] > nmax ? n[k] : nmax;
    }
    if (nseg <= 0 || nmax <= 0) return;
    const long blocks = ceil_div(nmax, 256) < 1024 ? ceil_div(nmax, 256) : 1024;
    hipLaunchKernelGGL(k_el_sum_segs, dim3(blocks, nseg), dim3(256), 0, s, S);
}

// dst[i] = src[i], i < n (a kernel rather than a copy call, so that a captured graph holds
// only kernel nodes)
void el_launch_copy(double* dst, const double* src, long n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_el_copy, dim3(ceil_div(n, 256)), dim3(256), 0, s, dst, src, n);
}
