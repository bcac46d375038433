// spme.hip -- smooth particle-mesh Ewald reciprocal space for point charges and point
// dipoles: energy, potential, field and force at every site.
//
// The reference's SPME (ewald_spme.f, spme_for.f, bspgen.f, bspcoe.f) spreads charges with
// cardinal B-splines of order p onto a K1 x K2 x K3 grid, transforms it, multiplies by the
// influence function B(m) C(m) and gathers potential derivatives back; this file does the
// same with the dipoles carried in the spread density (mu . grad of the spline product),
// the reciprocal part of the induced-dipole Ewald sum (ewald1p.f), to which it converges.
//
// Device steps (one stream, no host round trip), each over one or several grids:
//   k_sp_prep       per (site, axis): grid index, M_p, M_p', M_p'' (positions only: the
//                   solver's repeated dipole fields reuse them)
//   k_sp_tile       one workgroup per (plane, R rows) of the grid: the tiles of up to five
//                   densities in LDS as 64-bit fixed point (2^-44 e), contributions of every
//                   site whose support meets them added with LDS atomics (integer adds, so
//                   the grid does not depend on their order), then written as complex values
//   k_sp_fft x3     mixed-radix (2, 3, 5) Stockham passes in LDS along each axis; L lines
//                   of one axis per workgroup of 128 threads (L K <= 256 elements)
//   k_sp_influence  Q^(m) *= B(m) C(m); U = 1/2 sum B C |Q^|^2 as block partials
//                   (k_sp_influence_stress: also the six kernels B C akv2 k_a k_b that the
//                   per-site stress needs)
//   k_sp_fft x3     backward passes (unnormalised, conjugate twiddles)
//   k_sp_gather     one wavefront per site: phi, grad phi and (for forces) the Hessian of
//                   phi from the p^3 grid points, reduced over the lanes in a fixed order
//                   (k_sp_field: the field alone; k_sp_elec_final: every per-site
//                   reciprocal quantity of the polarizable electrostatics)
//   k_sp_usum       the energy partials in a fixed order
// Coulomb constant 1 (energy e^2/A); the caller scales.
#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "mbgpu_internal.h"

namespace {

constexpr double SP_FIX = 17592186044416.0;            // 2^44
constexpr double SP_UNFIX = 1.0 / 17592186044416.0;
constexpr double SP_PI = 3.141592653589793;

__device__ __forceinline__ double sp_wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int sp_wrap(long g, int K)
{
    long r = g % K;
    return (int)(r < 0 ? r + K : r);
}

template <int P, int NDER>
__device__ __forceinline__ void sp_splines_p(double w, double* a, double* d, double* d2)
{
    double m[P], m1[P], m2[P];
#pragma unroll
    for (int t = 0; t < P; ++t) m[t] = m1[t] = m2[t] = 0.0;
    m[0] = w;
    m[1] = 1.0 - w;
#pragma unroll
    for (int n = 3; n <= P; ++n) {
        if (n == P - 1)
#pragma unroll
            for (int t = 0; t < P; ++t) m2[t] = m[t];
        if (n == P)
#pragma unroll
            for (int t = 0; t < P; ++t) m1[t] = m[t];
        double prev = 0.0;
#pragma unroll
        for (int t = 0; t < n; ++t) {
            const double x = w + t;
            const double cur = t < n - 1 ? m[t] : 0.0;
            const double nx = (x * cur + ((double)n - x) * prev) / (double)(n - 1);
            prev = cur;
            m[t] = nx;
        }
    }
#pragma unroll
    for (int t = 0; t < P; ++t) {
        a[t] = m[t];
        if (NDER >= 1) d[t] = (t < P - 1 ? m1[t] : 0.0) - (t > 0 ? m1[t - 1] : 0.0);
        if (NDER >= 2)
            d2[t] = (t < P - 2 ? m2[t] : 0.0) - 2.0 * (t > 0 ? m2[t - 1] : 0.0) + (t > 1 ? m2[t - 2] : 0.0);
    }
}

// M_p(w + t) (and M_p', M_p'' when asked) for t = 0..p-1 by the recursion in the order
// (bspgen.f:27-100): M_n(x) = [x M_{n-1}(x) + (n - x) M_{n-1}(x - 1)] / (n - 1) from
// M_2(w) = w, M_2(w + 1) = 1 - w; M_n' = M_{n-1}(x) - M_{n-1}(x - 1).  Compiled per order
// so that the recursion stays in registers.
template <int NDER>
__device__ void sp_splines(int p, double w, double* a, double* d, double* d2)
{
    switch (p) {
    case 4: sp_splines_p<4, NDER>(w, a, d, d2); break;
    case 6: sp_splines_p<6, NDER>(w, a, d, d2); break;
    case 8: sp_splines_p<8, NDER>(w, a, d, d2); break;
    case 10: sp_splines_p<10, NDER>(w, a, d, d2); break;
    default: sp_splines_p<12, NDER>(w, a, d, d2); break;
    }
}

// scaled fractional coordinate u_a = K_a (a*_a . r); false for a position too far from the
// cell (or NaN) to index the grid
__device__ __forceinline__ bool sp_frac(const SpmeGeom& G, int a, double x, double y, double z, long& g0, double& w)
{
    const double u = G.K[a] * (G.rv[a][0] * x + G.rv[a][1] * y + G.rv[a][2] * z);
    if (!(fabs(u) < 1.0e9)) return false;
    const double fl = floor(u);
    g0 = (long)fl;
    w = u - fl;
    return true;
}

#define SP_MAX_SPREAD 5   // densities one spreading launch takes

__global__ __launch_bounds__(64) void k_sp_clear_flag(int* __restrict__ flag)
{
    if (threadIdx.x < 2) flag[threadIdx.x] = 0;
}

// Per (site, axis): the grid index of t = 0 (wrapped) and M_p, M_p', M_p'' at the p points;
// ok = 0 (and flag 1) for a position that cannot index the grid.  Layout: gi[i] = (g1, g2,
// g3, ok); th[i][a][k][t] (k: M, M', M'').
__global__ __launch_bounds__(256) void k_sp_prep(SpmeGeom G, int n, const double* __restrict__ x,
                                                 const double* __restrict__ y, const double* __restrict__ z,
                                                 int4* __restrict__ gi, double* __restrict__ th,
                                                 int* __restrict__ flag)
{
    __shared__ int okv[256];
    __shared__ int gv[256];
    const int t = blockIdx.x * 255 + threadIdx.x;   // (site, axis) pairs, 255 = 3 * 85 per block
    const int i = t / 3, a = t - 3 * (t / 3);
    const int p = G.p;
    const bool live = threadIdx.x < 255 && i < n;
    int good = 0, g = 0;
    if (live) {
        long g0 = 0;
        double w = 0.0;
        if (sp_frac(G, a, x[i], y[i], z[i], g0, w)) {
            good = 1;
            g = sp_wrap(g0, G.K[a]);
            double* T = th + ((long)i * 3 + a) * 3 * SPME_MAX_ORDER;
            sp_splines<2>(p, w, T, T + SPME_MAX_ORDER, T + 2 * SPME_MAX_ORDER);
        } else {
            flag[0] = 1;
        }
    }
    okv[threadIdx.x] = good;
    gv[threadIdx.x] = g;
    __syncthreads();
    if (live && a == 0) {
        const int b = threadIdx.x;
        gi[i] = make_int4(gv[b], gv[b + 1], gv[b + 2], okv[b] & okv[b + 1] & okv[b + 2]);
    }
}

// Mixed-radix Stockham transform of L lines of length K = 2^a 3^b 5^c along one axis.
// Line l of block (o, ib) starts at o * so + (ib * L + l) * si, element e at + e * se.
// Pass f with radix R after Ns = (product of the earlier radices): butterfly j < K/R reads
// elements j + q K/R (q < R), multiplies element q by w^(q k) with k = j mod Ns and
// w = exp(-2 pi i / (Ns R)) (conjugated backward), takes the length-R DFT and writes it to
// (j / Ns) Ns R + k + q Ns.  tw holds exp(-2 pi i t / K), t < K, staged in LDS with the
// lines; the radices are 4 (pairs of 2s), 2, 3, 5, each pass compiled for its radix.
__device__ __forceinline__ double2 sp_cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// v <- DFT_R(v) with exp(sgn 2 pi i q m / R), sgn = -1 forward, +1 backward
template <int R>
__device__ __forceinline__ void sp_dft(double sgn, double2 (&v)[R])
{
    if constexpr (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = make_double2(a.x + b.x, a.y + b.y);
        v[1] = make_double2(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 4) {
        // y1 = (v0 - v2) + W (v1 - v3), y3 = (v0 - v2) - W (v1 - v3), W = sgn i
        const double2 s02 = make_double2(v[0].x + v[2].x, v[0].y + v[2].y);
        const double2 d02 = make_double2(v[0].x - v[2].x, v[0].y - v[2].y);
        const double2 s13 = make_double2(v[1].x + v[3].x, v[1].y + v[3].y);
        const double2 d13 = make_double2(v[1].x - v[3].x, v[1].y - v[3].y);
        const double2 wd = make_double2(-sgn * d13.y, sgn * d13.x);
        v[0] = make_double2(s02.x + s13.x, s02.y + s13.y);
        v[2] = make_double2(s02.x - s13.x, s02.y - s13.y);
        v[1] = make_double2(d02.x + wd.x, d02.y + wd.y);
        v[3] = make_double2(d02.x - wd.x, d02.y - wd.y);
    } else if constexpr (R == 3) {
        const double h = 0.86602540378443864676;   // sin(2 pi / 3)
        const double2 t1 = make_double2(v[1].x + v[2].x, v[1].y + v[2].y);
        const double2 t2 = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
        const double2 t3 = make_double2(sgn * h * (v[1].x - v[2].x), sgn * h * (v[1].y - v[2].y));
        v[0] = make_double2(v[0].x + t1.x, v[0].y + t1.y);
        // y1 = t2 + i t3, y2 = t2 - i t3
        v[1] = make_double2(t2.x - t3.y, t2.y + t3.x);
        v[2] = make_double2(t2.x + t3.y, t2.y - t3.x);
    } else {   // R == 5
        const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;   // cos(2 pi/5), cos(4 pi/5)
        const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;    // sin(2 pi/5), sin(4 pi/5)
        const double2 t1 = make_double2(v[1].x + v[4].x, v[1].y + v[4].y);
        const double2 t2 = make_double2(v[2].x + v[3].x, v[2].y + v[3].y);
        const double2 t3 = make_double2(v[1].x - v[4].x, v[1].y - v[4].y);
        const double2 t4 = make_double2(v[2].x - v[3].x, v[2].y - v[3].y);
        const double2 a1 = make_double2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
        const double2 a2 = make_double2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
        const double2 b1 = make_double2(sgn * (s1 * t3.x + s2 * t4.x), sgn * (s1 * t3.y + s2 * t4.y));
        const double2 b2 = make_double2(sgn * (s2 * t3.x - s1 * t4.x), sgn * (s2 * t3.y - s1 * t4.y));
        v[0] = make_double2(v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y);
        v[1] = make_double2(a1.x - b1.y, a1.y + b1.x);   // a1 + i b1
        v[4] = make_double2(a1.x + b1.y, a1.y - b1.x);   // a1 - i b1
        v[2] = make_double2(a2.x - b2.y, a2.y + b2.x);
        v[3] = make_double2(a2.x + b2.y, a2.y - b2.x);
    }
}

// one pass of radix R over the L lines of in (K apart) into out; blockDim threads
template <int R>
__device__ __forceinline__ void sp_pass(const double2* __restrict__ in, double2* __restrict__ out, int K, int L,
                                        int Ns, const double2* __restrict__ tws, bool inverse)
{
    const int KR = K / R, ts = K / (Ns * R);
    const double sgn = inverse ? 1.0 : -1.0;
    for (int b = threadIdx.x; b < L * KR; b += blockDim.x) {
        const int l = b / KR, j = b - l * KR;
        const int k = j % Ns;
        const double2* li = in + l * K;
        double2 v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            v[q] = li[j + q * KR];
            if (q > 0 && k > 0) {
                double2 w = tws[q * k * ts];
                if (inverse) w.y = -w.y;
                v[q] = sp_cmul(v[q], w);
            }
        }
        sp_dft<R>(sgn, v);
        double2* lo = out + l * K + (j - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) lo[q * Ns] = v[q];
    }
}

// every pass of axis A over L lines staged in b0 (b1 the other buffer); returns the
// buffer that holds the result.  Ends with a barrier.
__device__ double2* sp_fft_lines(double2* b0, double2* b1, int K, int L, int nf, const int* fac,
                                 const double2* __restrict__ tws, bool inverse)
{
    double2* in = b0;
    double2* out = b1;
    int Ns = 1;
    for (int f = 0; f < nf; ++f) {
        const int R = fac[f];
        if (R == 4)
            sp_pass<4>(in, out, K, L, Ns, tws, inverse);
        else if (R == 2)
            sp_pass<2>(in, out, K, L, Ns, tws, inverse);
        else if (R == 3)
            sp_pass<3>(in, out, K, L, Ns, tws, inverse);
        else
            sp_pass<5>(in, out, K, L, Ns, tws, inverse);
        __syncthreads();
        double2* t = in;
        in = out;
        out = t;
        Ns *= R;
    }
    return in;
}

// line l / element e of the staged index idx: lines side by side for strided axes (so
// that neighbouring threads touch neighbouring lines), elements side by side otherwise
__device__ __forceinline__ void sp_line_elem(const SpmeAxis& A, int idx, int& l, int& e)
{
    if (A.se == 1) {
        l = idx / A.K;
        e = idx - l * A.K;
    } else {
        e = idx / A.L;
        l = idx - e * A.L;
    }
}

// blockIdx.y: the grid (grids gs points apart)
__global__ __launch_bounds__(SP_FFT_THREADS) void k_sp_fft(double2* __restrict__ grids, long gs, SpmeAxis A,
                                                           const double2* __restrict__ tw, int inverse,
                                                           const double* __restrict__ state)
{
    if (state && state[2] != 0.0) return;
    __shared__ double2 buf[2][SP_FFT_ELEMS > SPME_MAX_GRID ? SP_FFT_ELEMS : SPME_MAX_GRID];
    __shared__ double2 tws[SPME_MAX_GRID];
    double2* __restrict__ g = grids + blockIdx.y * gs;
    const int K = A.K, L = A.L;
    const int nib = A.ni / L;
    const int o = blockIdx.x / nib, ib = blockIdx.x - o * nib;
    const long base = (long)o * A.so + (long)ib * L * A.si;
    const int LK = L * K;
    for (int t = threadIdx.x; t < K; t += SP_FFT_THREADS) tws[t] = tw[t];
    for (int idx = threadIdx.x; idx < LK; idx += SP_FFT_THREADS) {
        int l, e;
        sp_line_elem(A, idx, l, e);
        buf[0][l * K + e] = g[base + (long)l * A.si + (long)e * A.se];
    }
    __syncthreads();
    const double2* res = sp_fft_lines(buf[0], buf[1], K, L, A.nf, A.fac, tws, inverse != 0);
    for (int idx = threadIdx.x; idx < LK; idx += SP_FFT_THREADS) {
        int l, e;
        sp_line_elem(A, idx, l, e);
        g[base + (long)l * A.si + (long)e * A.se] = res[l * K + e];
    }
}

// Sites by grid plane (gi.x), for the spreading tiles: bin_off[0..K1] offsets into
// bin_site, which lists the sites of each plane (their order within a plane is whatever
// the LDS counters give: the spread adds integers, so it does not depend on it).  Sites
// that cannot index the grid are left out.  One workgroup.
__global__ __launch_bounds__(1024) void k_sp_bin(int n, int K1, const int4* __restrict__ gi, int* __restrict__ bin_off,
                                                 int* __restrict__ bin_site)
{
    __shared__ int cnt[SPME_MAX_GRID], pos[SPME_MAX_GRID];
    const int tid = threadIdx.x;
    for (int t = tid; t < K1; t += 1024) cnt[t] = 0;
    __syncthreads();
    for (int i = tid; i < n; i += 1024) {
        const int4 g = gi[i];
        if (g.w) atomicAdd(&cnt[g.x], 1);
    }
    __syncthreads();
    if (tid == 0) {
        int a = 0;
        for (int t = 0; t < K1; ++t) {
            pos[t] = a;
            bin_off[t] = a;
            a += cnt[t];
        }
        bin_off[K1] = a;
    }
    __syncthreads();
    for (int i = tid; i < n; i += 1024) {
        const int4 g = gi[i];
        if (g.w) bin_site[atomicAdd(&pos[g.x], 1)] = i;
    }
}

// Spread by tiles: a workgroup owns plane g1 and rows r0..r0+R-1 of the grid (all g3) of
// each of the spec.ngr densities in LDS as 64-bit fixed point (2^-44 e), takes the sites
// of planes g1..g1+p-1 (the only ones whose support reaches g1) from the plane bins, keeps
// those whose rows meet the tile and adds their contributions with LDS atomics (integer
// adds: the tile does not depend on their order).  Sites i >= n are skipped.  Contribution
// of a site with charge c and dipole weights w_a = K_a (a*_a . mu) at (t1, t2, t3):
// c M1 M2 M3 + w1 M1' M2 M3 + w2 M1 M2' M3 + w3 M1 M2 M3'.  The tiles are written as
// complex values to grids + g * gs; with fwdz, after the forward transform of their rows
// (the axis-2 pass), from tw (the axis's twiddles) and A (its radices).
__global__ __launch_bounds__(256) void k_sp_tile(SpmeGeom G, int n, const int4* __restrict__ gi,
                                                 const double* __restrict__ th, const int* __restrict__ bin_off,
                                                 const int* __restrict__ bin_site, const double* __restrict__ q,
                                                 const double* __restrict__ mu, SpSpec spec,
                                                 double2* __restrict__ grids, long gs, int fwdz, SpmeAxis A,
                                                 const double2* __restrict__ tw, int* __restrict__ flag,
                                                 const double* __restrict__ state)
{
    if (state && state[2] != 0.0) return;
    extern __shared__ unsigned long long tile[];   // [ngr][R][K3], then (fwdz) 2 x [R][K3] + [K3] complex
    constexpr int CU = 4;                          // candidates per thread and round
    __shared__ int list[256 * CU], lt1[256 * CU];
    __shared__ int cnt, boff[SPME_MAX_ORDER + 1], bstart[SPME_MAX_ORDER], bcnt[SPME_MAX_ORDER];
    const int K1 = G.K[0], K2 = G.K[1], K3 = G.K[2], p = G.p;
    const int R = G.rows, nrb = K2 / R, ngr = spec.ngr;
    const int TS = R * K3;
    const int g1 = blockIdx.x / nrb, r0 = (blockIdx.x - g1 * nrb) * R;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < ngr * TS; idx += 256) tile[idx] = 0ull;
    if (tid < p) {
        // the p plane bins g1 + t1, read side by side, then as one candidate range [0, boff[p])
        int b = g1 + tid;
        b -= b >= K1 ? K1 : 0;
        const int o0 = bin_off[b], o1 = bin_off[b + 1];
        bstart[tid] = o0;
        bcnt[tid] = o1 - o0;
    }
    __syncthreads();
    if (tid == 0) {
        int a = 0;
        for (int t1 = 0; t1 < p; ++t1) {
            boff[t1] = a;
            a += bcnt[t1];
        }
        boff[p] = a;
    }
    __syncthreads();
    const int ncand = boff[p];
    double msc = 1.0;
    if (spec.scale_state > 0 && state) {
        const double v = state[spec.scale_state];
        msc = v > 0.0 ? v : 1.0;
    }
    bool big = false;
    for (int base = 0; base < ncand; base += 256 * CU) {
        if (tid == 0) cnt = 0;
        __syncthreads();
        // CU candidates per thread, their loads issued together
        int ci[CU], ct[CU];
#pragma unroll
        for (int u = 0; u < CU; ++u) {
            const int c = base + tid + 256 * u;
            ci[u] = n;
            ct[u] = 0;
            if (c < ncand) {
                int t1 = 0;
                while (c >= boff[t1 + 1]) ++t1;
                ct[u] = t1;
                ci[u] = bin_site[bstart[t1] + (c - boff[t1])];
            }
        }
        int gy[CU];
#pragma unroll
        for (int u = 0; u < CU; ++u) gy[u] = ci[u] < n ? gi[ci[u]].y : -1;
#pragma unroll
        for (int u = 0; u < CU; ++u) {
            if (gy[u] < 0) continue;
            int d = gy[u] - r0;
            d += d < 0 ? K2 : 0;
            bool hit = false;
            for (int t2 = 0; t2 < p; ++t2) {
                int r = d - t2;
                r += r < 0 ? K2 : 0;
                hit |= r < R;
            }
            if (hit) {
                const int k = atomicAdd(&cnt, 1);
                list[k] = ci[u];
                lt1[k] = ct[u];
            }
        }
        __syncthreads();
        const int m = cnt;
        for (int wk = tid; wk < m * p; wk += 256) {
            const int e = wk / p, t2 = wk - e * p;
            const int i2 = list[e], t1 = lt1[e];
            const int4 g = gi[i2];
            int r = g.y - r0 - t2;
            r += r < 0 ? K2 : 0;
            r += r < 0 ? K2 : 0;
            if (r >= R) continue;
            const double* T = th + (long)i2 * 9 * SPME_MAX_ORDER;
            const double a1 = T[t1], d1 = T[SPME_MAX_ORDER + t1];
            const double a2 = T[3 * SPME_MAX_ORDER + t2], d2 = T[4 * SPME_MAX_ORDER + t2];
            const double* T3 = T + 6 * SPME_MAX_ORDER;
            const double qi = q ? q[i2] : 0.0;
            double mv[3] = {0.0, 0.0, 0.0};
            if (mu)
                for (int cc = 0; cc < 3; ++cc) mv[cc] = mu[3 * i2 + cc] * msc;
            double wa[3];
            for (int a = 0; a < 3; ++a) wa[a] = G.K[a] * (G.rv[a][0] * mv[0] + G.rv[a][1] * mv[1] + G.rv[a][2] * mv[2]);
            for (int gr = 0; gr < ngr; ++gr) {
                const double cg = (spec.q_on[gr] ? qi : 0.0) + (spec.comp[gr] >= 0 ? mv[spec.comp[gr]] : 0.0);
                const double mo = spec.mu_on[gr] ? 1.0 : 0.0;
                const double c12 = cg * a1 * a2 + mo * (wa[0] * d1 * a2 + wa[1] * a1 * d2);
                const double w3 = mo * wa[2] * a1 * a2;
                unsigned long long* row = tile + gr * TS + r * K3;
                for (int t3 = 0; t3 < p; ++t3) {
                    const double v = c12 * T3[t3] + w3 * T3[SPME_MAX_ORDER + t3];
                    big |= !(fabs(v) < 1024.0);
                    int g3 = g.z - t3;
                    g3 += g3 < 0 ? K3 : 0;
                    atomicAdd(row + g3, (unsigned long long)llrint(v * SP_FIX));
                }
            }
        }
        __syncthreads();
    }
    if (big) flag[0] = 2;
    if (!fwdz) {
        for (int idx = tid; idx < ngr * TS; idx += 256) {
            const int gr = idx / TS, rem = idx - gr * TS;
            const int r = rem / K3, c3 = rem - r * K3;
            grids[gr * gs + ((long)g1 * K2 + r0 + r) * K3 + c3] =
                make_double2((double)(long long)tile[idx] * SP_UNFIX, 0.0);
        }
        return;
    }
    double2* b0 = reinterpret_cast<double2*>(tile + ngr * TS);
    double2* b1 = b0 + TS;
    double2* tws = b1 + TS;
    for (int t = tid; t < K3; t += 256) tws[t] = tw[t];
    for (int gr = 0; gr < ngr; ++gr) {
        for (int idx = tid; idx < TS; idx += 256) b0[idx] = make_double2((double)(long long)tile[gr * TS + idx] * SP_UNFIX, 0.0);
        __syncthreads();
        const double2* res = sp_fft_lines(b0, b1, K3, R, A.nf, A.fac, tws, false);
        double2* out = grids + gr * gs + ((long)g1 * K2 + r0) * K3;
        for (int idx = tid; idx < TS; idx += 256) out[idx] = res[idx];
        __syncthreads();
    }
}

// B(m) C(m) at flat index i of the grid, C(m) = exp(-pi^2 m^2 / alpha^2) / (pi V m^2),
// m = sum_a m_a a*_a with m_a folded into (-K_a/2, K_a/2]; the Cartesian m into mv and m^2
// into msq (0 at m = 0)
__device__ __forceinline__ double sp_bc(const SpmeGeom& G, const double* __restrict__ bmod, long i, double* mv,
                                        double& msq)
{
    const int K1 = G.K[0], K2 = G.K[1], K3 = G.K[2];
    const int m3 = (int)(i % K3), m2 = (int)((i / K3) % K2), m1 = (int)(i / ((long)K2 * K3));
    const int a1 = m1 <= K1 / 2 ? m1 : m1 - K1, a2 = m2 <= K2 / 2 ? m2 : m2 - K2, a3 = m3 <= K3 / 2 ? m3 : m3 - K3;
    msq = 0.0;
    for (int d = 0; d < 3; ++d) mv[d] = 0.0;
    if (!(a1 || a2 || a3)) return 0.0;
    for (int d = 0; d < 3; ++d) mv[d] = a1 * G.rv[0][d] + a2 * G.rv[1][d] + a3 * G.rv[2][d];
    msq = mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2];
    return bmod[m1] * bmod[K1 + m2] * bmod[K1 + K2 + m3] * exp(-SP_PI * SP_PI * msq / (G.alpha * G.alpha)) /
           (SP_PI * G.vol * msq);
}

// Q^(m) *= B(m) C(m) on grid blockIdx.y; with part, block partials of 1/2 B C |Q^|^2 into
// part[blockIdx.y * gridDim.x + blockIdx.x]
__global__ __launch_bounds__(256) void k_sp_influence(SpmeGeom G, double2* __restrict__ grids, long gs,
                                                      const double* __restrict__ bmod, double* __restrict__ part,
                                                      const double* __restrict__ state)
{
    if (state && state[2] != 0.0) return;
    __shared__ double red[4];
    double2* __restrict__ g = grids + blockIdx.y * gs;
    const long ng = (long)G.K[0] * G.K[1] * G.K[2];
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    if (i < ng) {
        double mv[3], msq;
        const double bc = sp_bc(G, bmod, i, mv, msq);
        double2 v = g[i];
        e = 0.5 * bc * (v.x * v.x + v.y * v.y);
        v.x *= bc;
        v.y *= bc;
        g[i] = v;
    }
    if (!part) return;
    e = sp_wsum(e);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) part[(long)blockIdx.y * gridDim.x + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// src *= B C in place, and dst[t] = B C akv2 k_a k_b src (k = 2 pi m, akv2 = 2 (1 / k^2 +
// 1 / (4 alpha^2)); t = xx xy xz yy yz zz), dst grids gs points apart
__global__ __launch_bounds__(256) void k_sp_influence_stress(SpmeGeom G, double2* __restrict__ src,
                                                             double2* __restrict__ dst, long gs,
                                                             const double* __restrict__ bmod)
{
    const long ng = (long)G.K[0] * G.K[1] * G.K[2];
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= ng) return;
    double mv[3], msq;
    const double bc = sp_bc(G, bmod, i, mv, msq);
    const double2 v = src[i];
    src[i] = make_double2(bc * v.x, bc * v.y);
    const double k2 = 4.0 * SP_PI * SP_PI * msq;
    const double f = msq > 0.0 ? bc * 2.0 * (1.0 / k2 + 0.25 / (G.alpha * G.alpha)) * 4.0 * SP_PI * SP_PI : 0.0;
    const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
    for (int t = 0; t < 6; ++t) {
        const double w = f * mv[ia[t]] * mv[ib[t]];
        dst[t * gs + i] = make_double2(w * v.x, w * v.y);
    }
}

// Axis 0 of grid blockIdx.y as one step: the forward pass, Q^ *= B C, the backward pass
// (the last forward and the first backward pass of the convolution, with the influence
// function between them, without the energy)
__global__ __launch_bounds__(SP_FFT_THREADS) void k_sp_conv0(SpmeGeom G, double2* __restrict__ grids, long gs,
                                                             SpmeAxis A, const double2* __restrict__ tw,
                                                             const double* __restrict__ bmod,
                                                             const double* __restrict__ state)
{
    if (state && state[2] != 0.0) return;
    __shared__ double2 buf[2][SP_FFT_ELEMS > SPME_MAX_GRID ? SP_FFT_ELEMS : SPME_MAX_GRID];
    __shared__ double2 tws[SPME_MAX_GRID];
    double2* __restrict__ g = grids + blockIdx.y * gs;
    const int K = A.K, L = A.L;
    const long base = (long)blockIdx.x * L;   // axis 0: lines are the K2 K3 columns, side by side
    const int LK = L * K;
    for (int t = threadIdx.x; t < K; t += SP_FFT_THREADS) tws[t] = tw[t];
    for (int idx = threadIdx.x; idx < LK; idx += SP_FFT_THREADS) {
        int l, e;
        sp_line_elem(A, idx, l, e);
        buf[0][l * K + e] = g[base + l + (long)e * A.se];
    }
    __syncthreads();
    double2* res = sp_fft_lines(buf[0], buf[1], K, L, A.nf, A.fac, tws, false);
    for (int idx = threadIdx.x; idx < LK; idx += SP_FFT_THREADS) {
        const int l = idx / K, e = idx - l * K;
        double mv[3], msq;
        const double bc = sp_bc(G, bmod, (long)e * A.se + base + l, mv, msq);
        double2 v = res[idx];
        res[idx] = make_double2(bc * v.x, bc * v.y);
    }
    __syncthreads();
    const double2* out = sp_fft_lines(res, res == buf[0] ? buf[1] : buf[0], K, L, A.nf, A.fac, tws, true);
    for (int idx = threadIdx.x; idx < LK; idx += SP_FFT_THREADS) {
        int l, e;
        sp_line_elem(A, idx, l, e);
        g[base + l + (long)e * A.se] = out[l * K + e];
    }
}

// energy of grid blockIdx.x from its nb partials
__global__ __launch_bounds__(256) void k_sp_usum(int nb, const double* __restrict__ part, double* __restrict__ out)
{
    __shared__ double red[4];
    const double* pp = part + (long)blockIdx.x * nb;
    double a = 0.0;
    for (int k = threadIdx.x; k < nb; k += 256) a += pp[k];
    a = sp_wsum(a);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// d/dr_d = sum_a K_a a*_a[d] d/du_a
__device__ __forceinline__ void sp_ga(const SpmeGeom& G, double (&ga)[3][3])
{
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) ga[a][d] = G.K[a] * G.rv[a][d];
}

// Cartesian gradient from the u-derivatives s[0..2]
__device__ __forceinline__ void sp_cart_grad(const double (&ga)[3][3], const double* s, double* gr)
{
    for (int d = 0; d < 3; ++d) gr[d] = ga[0][d] * s[0] + ga[1][d] * s[1] + ga[2][d] * s[2];
}

// Cartesian Hessian times m from the u-second derivatives h = (00, 11, 22, 01, 02, 12)
__device__ __forceinline__ void sp_cart_hess_mul(const double (&ga)[3][3], const double* h, const double* m,
                                                 double* out)
{
    const double hu[3][3] = {{h[0], h[3], h[4]}, {h[3], h[1], h[5]}, {h[4], h[5], h[2]}};
    for (int d = 0; d < 3; ++d) {
        double hm = 0.0;
        for (int e = 0; e < 3; ++e) {
            double hh = 0.0;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) hh += ga[a][d] * ga[b][e] * hu[a][b];
            hm += hh * m[e];
        }
        out[d] = hm;
    }
}

// Sums over the p^3 grid points of one site (lanes in turn, then a fixed-order wavefront
// reduction): NV = 1 the value; 4 value and u-gradient; 10 also the u-Hessian (00 11 22 01
// 02 12); 9 u-gradient and u-Hessian without the value.  Grid points index g[...].x.
template <int NV>
__device__ __forceinline__ void sp_site_sums(const SpmeGeom& G, int4 gg, const double* __restrict__ T,
                                             const double2* __restrict__ g, double* s)
{
    constexpr int MO = SPME_MAX_ORDER;
    const int p = G.p, K1 = G.K[0], K2 = G.K[1], K3 = G.K[2];
    const int lane = threadIdx.x & 63;
    double a[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) a[k] = 0.0;
    const int np3 = p * p * p;
    for (int fidx = lane; fidx < np3; fidx += 64) {
        const int t1 = fidx / (p * p), t2 = (fidx / p) % p, t3 = fidx % p;
        int g1 = gg.x - t1, g2 = gg.y - t2, g3 = gg.z - t3;
        g1 += g1 < 0 ? K1 : 0;
        g2 += g2 < 0 ? K2 : 0;
        g3 += g3 < 0 ? K3 : 0;
        const double v = g[((long)g1 * K2 + g2) * K3 + g3].x;
        const double A1 = T[t1], A2 = T[3 * MO + t2], A3 = T[6 * MO + t3];
        if constexpr (NV == 1) {
            a[0] += v * A1 * A2 * A3;
        } else {
            const double D1 = T[MO + t1], D2 = T[4 * MO + t2], D3 = T[7 * MO + t3];
            constexpr int o = NV == 9 ? -1 : 0;
            if constexpr (NV != 9) a[0] += v * A1 * A2 * A3;
            a[o + 1] += v * D1 * A2 * A3;
            a[o + 2] += v * A1 * D2 * A3;
            a[o + 3] += v * A1 * A2 * D3;
            if constexpr (NV >= 9) {
                const double E1 = T[2 * MO + t1], E2 = T[5 * MO + t2], E3 = T[8 * MO + t3];
                a[o + 4] += v * E1 * A2 * A3;
                a[o + 5] += v * A1 * E2 * A3;
                a[o + 6] += v * A1 * A2 * E3;
                a[o + 7] += v * D1 * D2 * A3;
                a[o + 8] += v * D1 * A2 * D3;
                a[o + 9] += v * A1 * D2 * D3;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] = sp_wsum(a[k]);
}

// one wavefront per site (4 per block): phi, the field and (HESS, forces asked for) the
// force -q grad phi - (grad grad phi) mu
template <bool HESS>
__global__ __launch_bounds__(256) void k_sp_gather(SpmeGeom G, int n, const int4* __restrict__ gi,
                                                   const double* __restrict__ th, const double* __restrict__ q,
                                                   const double* __restrict__ mu, const double2* __restrict__ g,
                                                   double* __restrict__ phi, double* __restrict__ ef,
                                                   double* __restrict__ f)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const int4 gg = gi[i];
    if (!gg.w) return;
    const double* T = th + (long)i * 9 * SPME_MAX_ORDER;
    double s[10];
    if (HESS)
        sp_site_sums<10>(G, gg, T, g, s);
    else
        sp_site_sums<4>(G, gg, T, g, s);
    if (lane != 0) return;
    double ga[3][3], gr[3];
    sp_ga(G, ga);
    sp_cart_grad(ga, s + 1, gr);
    if (phi) phi[i] = s[0];
    if (ef)
        for (int d = 0; d < 3; ++d) ef[3 * i + d] = -gr[d];
    if (HESS && f) {
        const double qi = q[i];
        double m[3] = {0.0, 0.0, 0.0};
        if (mu)
            for (int d = 0; d < 3; ++d) m[d] = mu[3 * i + d];
        double hm[3];
        sp_cart_hess_mul(ga, s + 4, m, hm);
        for (int d = 0; d < 3; ++d) f[3 * i + d] = -qi * gr[d] - hm[d];
    }
}

// E = -grad phi of grid blockIdx.y at site i, times fac, into out[(3 y + c) * os + i]
// (zeros for a site that cannot index the grid); one wavefront per site, 4 per block
__global__ __launch_bounds__(256) void k_sp_field(SpmeGeom G, int n, const int4* __restrict__ gi,
                                                  const double* __restrict__ th, const double2* __restrict__ grids,
                                                  long gs, double fac, double* __restrict__ out, long os,
                                                  const double* __restrict__ state, int scale_state)
{
    if (state && state[2] != 0.0) return;
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    if (scale_state > 0 && state) {
        // the spreading's power-of-two scale of the dipoles, divided out exactly
        const double v = state[scale_state];
        if (v > 0.0) fac /= v;
    }
    const int4 gg = gi[i];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if (gg.w) sp_site_sums<4>(G, gg, th + (long)i * 9 * SPME_MAX_ORDER, grids + blockIdx.y * gs, s);
    if (lane != 0) return;
    double ga[3][3], gr[3];
    sp_ga(G, ga);
    sp_cart_grad(ga, s + 1, gr);
    double* o = out + 3L * blockIdx.y * os + i;
    for (int c = 0; c < 3; ++c) o[c * os] = -fac * gr[c];
}

// Every per-site reciprocal quantity of the polarizable electrostatics (the quantities the
// Ewald k-sum's final site pass accumulates, ewald1p.f:550-693, 771-825) from the
// transformed grids of one evaluation:
//   grid 0      B C S^ of the density q + mu: phi, grad phi, grad grad phi
//   grid 1      B C Q^ of the charges alone: phi_q
//   grids 2..4  B C M_b^ of dipole component b as a charge density: grad and grad grad Lambda_b
//   grids 5..10 B C akv2 k_a k_b S^ (xx xy xz yy yz zz): Psi_ab and grad Psi_ab
// With kap the factor from these full-space convolutions to the half-space k-sums
// (V / 8 pi), site j gets, as the k-sum pass's partial sums:
//   [0..2]  kap (-q grad phi - (grad grad phi) mu)      the force
//   [3]     kap phi_q,  [4] kap (phi - phi_q),  [5] q [3]
//   [6 + 3a + b]  kap (delta_ab D - G_ab + P_ab + mu_b d_a phi), D = q phi + mu . grad phi,
//           G_ab = q Psi_ab + mu . grad Psi_ab, P_ab = -(q d_a Lambda_b + (grad grad Lambda_b mu)_a)
// into ks[t * n + j], and block partials [nb][8] of: q [3], kap D, and the six k-sum stress
// components (xx yy zz xy xz yz) of the site, [6 + 3a + b] - delta_ab kap D.
__global__ __launch_bounds__(256) void k_sp_elec_final(SpmeGeom G, int n, const int4* __restrict__ gi,
                                                       const double* __restrict__ th,
                                                       const double* __restrict__ q, const double* __restrict__ mu,
                                                       const double2* __restrict__ grids, long gs, double kap,
                                                       double* __restrict__ ks, double* __restrict__ blk)
{
    __shared__ double red[4][8];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = blockIdx.x * 4 + wv;
    double out[15], kd = 0.0;
    for (int t = 0; t < 15; ++t) out[t] = 0.0;
    if (i < n) {
        const int4 gg = gi[i];
        if (gg.w) {
            const double* T = th + (long)i * 9 * SPME_MAX_ORDER;
            double ga[3][3];
            sp_ga(G, ga);
            const double qi = q[i];
            const double m[3] = {mu[3 * i], mu[3 * i + 1], mu[3 * i + 2]};
            double s[10], gr[3], hm[3];
            sp_site_sums<10>(G, gg, T, grids, s);
            const double phi = s[0];
            sp_cart_grad(ga, s + 1, gr);
            sp_cart_hess_mul(ga, s + 4, m, hm);
            for (int d = 0; d < 3; ++d) out[d] = kap * (-qi * gr[d] - hm[d]);
            const double D = qi * phi + m[0] * gr[0] + m[1] * gr[1] + m[2] * gr[2];
            double st[9];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) st[3 * a + b] = (a == b ? D : 0.0) + m[b] * gr[a];
            sp_site_sums<1>(G, gg, T, grids + gs, s);
            out[3] = kap * s[0];
            out[4] = kap * (phi - s[0]);
            out[5] = qi * out[3];
            for (int b = 0; b < 3; ++b) {
                double lg[3], lh[3];
                sp_site_sums<9>(G, gg, T, grids + (2 + b) * gs, s);
                sp_cart_grad(ga, s, lg);
                sp_cart_hess_mul(ga, s + 3, m, lh);
                for (int a = 0; a < 3; ++a) st[3 * a + b] -= qi * lg[a] + lh[a];
            }
            const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
            for (int t = 0; t < 6; ++t) {
                double pg[3];
                sp_site_sums<4>(G, gg, T, grids + (5 + t) * gs, s);
                sp_cart_grad(ga, s + 1, pg);
                const double Gt = qi * s[0] + m[0] * pg[0] + m[1] * pg[1] + m[2] * pg[2];
                st[3 * ia[t] + ib[t]] -= Gt;
                if (ia[t] != ib[t]) st[3 * ib[t] + ia[t]] -= Gt;
            }
            for (int t = 0; t < 9; ++t) out[6 + t] = kap * st[t];
            kd = kap * D;
        }
        if (lane == 0)
            for (int t = 0; t < 15; ++t) ks[(long)t * n + i] = out[t];
    }
    if (lane == 0) {
        red[wv][0] = out[5];
        red[wv][1] = kd;
        red[wv][2] = out[6] - kd;
        red[wv][3] = out[10] - kd;
        red[wv][4] = out[14] - kd;
        red[wv][5] = out[7];
        red[wv][6] = out[8];
        red[wv][7] = out[11];
    }
    __syncthreads();
    if (threadIdx.x < 8)
        blk[8L * blockIdx.x + threadIdx.x] =
            ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

}  // namespace

// ---------------------------------------------------------------------- launchers ----

void sp_launch_prep(SpmeBuffers& B, const SpmeGeom& G, int n, const double* x, const double* y, const double* z,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_sp_clear_flag, dim3(1), dim3(64), 0, s, B.flag);
    if (n > 0)
        hipLaunchKernelGGL(k_sp_prep, dim3((unsigned)((3L * n + 254) / 255)), dim3(256), 0, s, G, n, x, y, z, B.gi,
                           B.th, B.flag);
    hipLaunchKernelGGL(k_sp_bin, dim3(1), dim3(1024), 0, s, n, G.K[0], B.gi, B.bin_off, B.bin_site);
}

size_t sp_spread_lds(const SpmeGeom& G, int ngr, bool fwdz)
{
    const size_t ts = (size_t)G.rows * G.K[2];
    return sizeof(unsigned long long) * ngr * ts + (fwdz ? sizeof(double2) * (2 * ts + G.K[2]) : 0);
}

void sp_launch_spread(SpmeBuffers& B, const SpmeGeom& G, int n, const double* q, const double* mu, const SpSpec& spec,
                      double2* grids, bool fwdz, const double* state, hipStream_t s)
{
    hipLaunchKernelGGL(k_sp_tile, dim3((unsigned)(G.K[0] * (G.K[1] / G.rows))), dim3(256),
                       sp_spread_lds(G, spec.ngr, fwdz), s, G, n, B.gi, B.th, B.bin_off, B.bin_site, q, mu, spec,
                       grids, B.ng_cap, fwdz ? 1 : 0, B.axis[2], B.tw + B.twoff[2], B.flag, state);
}

void sp_launch_fft(SpmeBuffers& B, const SpmeGeom& G, double2* grids, int ngr, bool inverse, const double* state,
                   hipStream_t s, bool zdone)
{
    (void)G;
    // axes in the order 2, 1, 0 forward (2 skipped when the spreading did it), 0, 1, 2 back
    for (int k = 0; k < 3; ++k) {
        const int a = inverse ? k : 2 - k;
        if (!inverse && a == 2 && zdone) continue;
        const SpmeAxis& A = B.axis[a];
        const unsigned nblk = (unsigned)((long)A.no * (A.ni / A.L));
        hipLaunchKernelGGL(k_sp_fft, dim3(nblk, (unsigned)ngr), dim3(SP_FFT_THREADS), 0, s, grids, B.ng_cap, A,
                           B.tw + B.twoff[a], inverse ? 1 : 0, state);
    }
}

void sp_launch_convolve(SpmeBuffers& B, const SpmeGeom& G, double2* grids, int ngr, bool zdone, const double* state,
                        hipStream_t s)
{
    // forward axes 2 (unless the spreading did it) and 1, axis 0 forward - influence -
    // backward in one step, backward 1, 2
    for (int a : {2, 1}) {
        if (a == 2 && zdone) continue;
        const SpmeAxis& A = B.axis[a];
        hipLaunchKernelGGL(k_sp_fft, dim3((unsigned)((long)A.no * (A.ni / A.L)), (unsigned)ngr), dim3(SP_FFT_THREADS),
                           0, s, grids, B.ng_cap, A, B.tw + B.twoff[a], 0, state);
    }
    const SpmeAxis& A0 = B.axis[0];
    hipLaunchKernelGGL(k_sp_conv0, dim3((unsigned)(A0.ni / A0.L), (unsigned)ngr), dim3(SP_FFT_THREADS), 0, s, G, grids,
                       B.ng_cap, A0, B.tw + B.twoff[0], B.bmod, state);
    for (int a : {1, 2}) {
        const SpmeAxis& A = B.axis[a];
        hipLaunchKernelGGL(k_sp_fft, dim3((unsigned)((long)A.no * (A.ni / A.L)), (unsigned)ngr), dim3(SP_FFT_THREADS),
                           0, s, grids, B.ng_cap, A, B.tw + B.twoff[a], 1, state);
    }
}

void sp_launch_influence(SpmeBuffers& B, const SpmeGeom& G, double2* grids, int ngr, bool energy,
                         const double* state, hipStream_t s)
{
    const int nb = (int)((spme_points(G) + 255) / 256);
    hipLaunchKernelGGL(k_sp_influence, dim3((unsigned)nb, (unsigned)ngr), dim3(256), 0, s, G, grids, B.ng_cap, B.bmod,
                       energy ? B.part : nullptr, state);
}

void sp_launch_usum(SpmeBuffers& B, const SpmeGeom& G, int ngr, hipStream_t s)
{
    const int nb = (int)((spme_points(G) + 255) / 256);
    hipLaunchKernelGGL(k_sp_usum, dim3((unsigned)ngr), dim3(256), 0, s, nb, B.part, B.out);
}

void sp_launch_influence_stress(SpmeBuffers& B, const SpmeGeom& G, double2* src, double2* dst, hipStream_t s)
{
    const int nb = (int)((spme_points(G) + 255) / 256);
    hipLaunchKernelGGL(k_sp_influence_stress, dim3((unsigned)nb), dim3(256), 0, s, G, src, dst, B.ng_cap, B.bmod);
}

void sp_launch_field(SpmeBuffers& B, const SpmeGeom& G, int n, const double2* grids, int ngr, double fac,
                     double* out, long ostride, const double* state, hipStream_t s, int scale_state)
{
    if (n > 0)
        hipLaunchKernelGGL(k_sp_field, dim3((unsigned)((n + 3) / 4), (unsigned)ngr), dim3(256), 0, s, G, n, B.gi, B.th,
                           grids, B.ng_cap, fac, out, ostride, state, scale_state);
}

void sp_launch_elec_final(SpmeBuffers& B, const SpmeGeom& G, int n, const double* q, const double* mu,
                          const double2* grids, double kap, double* ks, double* blk, hipStream_t s)
{
    if (n > 0)
        hipLaunchKernelGGL(k_sp_elec_final, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, G, n, B.gi, B.th, q, mu,
                           grids, B.ng_cap, kap, ks, blk);
}

int spme_launch(SpmeBuffers& B, const SpmeGeom& G, int n, const double* x, const double* y, const double* z,
                const double* q, const double* mu, double* phi, double* ef, double* f, hipStream_t s)
{
    sp_launch_prep(B, G, n, x, y, z, s);
    SpSpec spec{};
    spec.ngr = 1;
    spec.q_on[0] = 1;
    spec.mu_on[0] = mu ? 1 : 0;
    spec.comp[0] = -1;
    sp_launch_spread(B, G, n, q, mu, spec, B.grid, true, nullptr, s);
    sp_launch_fft(B, G, B.grid, 1, false, nullptr, s, true);
    sp_launch_influence(B, G, B.grid, 1, true, nullptr, s);
    sp_launch_usum(B, G, 1, s);
    sp_launch_fft(B, G, B.grid, 1, true, nullptr, s, false);
    if (n > 0 && f)
        hipLaunchKernelGGL(k_sp_gather<true>, dim3((n + 3) / 4), dim3(256), 0, s, G, n, B.gi, B.th, q, mu, B.grid, phi,
                           ef, f);
    else if (n > 0 && (phi || ef))
        hipLaunchKernelGGL(k_sp_gather<false>, dim3((n + 3) / 4), dim3(256), 0, s, G, n, B.gi, B.th, q, mu, B.grid, phi,
                           ef, f);
    return (int)hipGetLastError();
}

// --------------------------------------------------------------------------- plan ----

namespace {

bool sp_smooth(int K)
{
    for (int r : {2, 3, 5})
        while (K % r == 0) K /= r;
    return K == 1;
}

int sp_largest_divisor(int n, int cap)
{
    for (int d = std::min(n, std::max(1, cap)); d > 1; --d)
        if (n % d == 0) return d;
    return 1;
}

// radices of the transform passes (the 2s in pairs as 4s, then a lone 2, the 3s, the 5s)
// and L, the lines per workgroup: the largest divisor of ni with L K <= SP_FFT_ELEMS
void sp_axis_plan(SpmeAxis& A)
{
    int K = A.K, nf = 0;
    for (int r : {4, 2, 3, 5})
        while (K % r == 0) {
            A.fac[nf++] = r;
            K /= r;
        }
    A.nf = nf;
    A.L = sp_largest_divisor(A.ni, SP_FFT_ELEMS / A.K);
}

template <typename T>
void sp_free(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

#define SPCHK(x)                                                                 \
    do {                                                                         \
        hipError_t e_ = (x);                                                     \
        if (e_ != hipSuccess) {                                                  \
            err = std::string(#x) + ": " + hipGetErrorString(e_);                \
            return -1;                                                           \
        }                                                                        \
    } while (0)

}  // namespace

void spme_free(SpmeBuffers& B)
{
    sp_free(B.gi);
    sp_free(B.th);
    sp_free(B.bin_off);
    sp_free(B.bin_site);
    sp_free(B.grid);
    sp_free(B.bmod);
    sp_free(B.tw);
    sp_free(B.part);
    sp_free(B.out);
    sp_free(B.flag);
    B.n_cap = 0;
    B.ng_cap = 0;
    B.ngr_cap = 0;
    B.cfg[0] = 0;
}

int spme_plan(SpmeBuffers& B, SpmeGeom& G, const double cell[9], int order, const int grid[3], double alpha, int n,
              int ngr, std::string& err)
{
    char buf[200];
    if (!cell || !grid) {
        err = "null cell or grid";
        return -1;
    }
    if (order < 4 || order > SPME_MAX_ORDER || (order & 1)) {
        snprintf(buf, sizeof buf, "order must be even, 4..%d (got %d)", SPME_MAX_ORDER, order);
        err = buf;
        return -1;
    }
    for (int a = 0; a < 3; ++a) {
        const int K = grid[a];
        if (K < 8 || K > SPME_MAX_GRID || !sp_smooth(K) || K < order) {
            snprintf(buf, sizeof buf, "grid[%d] = %d: each grid size must be 2^a 3^b 5^c in [max(8, order), %d]", a,
                     K, SPME_MAX_GRID);
            err = buf;
            return -1;
        }
    }
    if (!(alpha > 0.0)) {
        err = "alpha must be positive";
        return -1;
    }
    if (n < 0 || ngr < 1 || ngr > SP_MAX_GRIDS) {
        err = "site or grid count out of range";
        return -1;
    }
    G = SpmeGeom{};
    {
        const double *a = cell, *b = cell + 3, *cc = cell + 6;
        const double bxc[3] = {b[1] * cc[2] - b[2] * cc[1], b[2] * cc[0] - b[0] * cc[2], b[0] * cc[1] - b[1] * cc[0]};
        const double cxa[3] = {cc[1] * a[2] - cc[2] * a[1], cc[2] * a[0] - cc[0] * a[2], cc[0] * a[1] - cc[1] * a[0]};
        const double axb[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        const double v = a[0] * bxc[0] + a[1] * bxc[1] + a[2] * bxc[2];
        if (!(fabs(v) > 0.0) || !std::isfinite(v)) {
            err = "the cell has no volume";
            return -1;
        }
        for (int d = 0; d < 3; ++d) {
            G.rv[0][d] = bxc[d] / v;
            G.rv[1][d] = cxa[d] / v;
            G.rv[2][d] = axb[d] / v;
        }
        G.vol = fabs(v);
    }
    G.alpha = alpha;
    G.p = order;
    for (int a = 0; a < 3; ++a) G.K[a] = grid[a];
    const int K1 = grid[0], K2 = grid[1], K3 = grid[2];
    // tile rows: LDS of up to SP_MAX_SPREAD fixed-point tiles, the two complex row buffers
    // of the fused axis-2 pass and its twiddles within 60 000 bytes (the site lists take
    // 2 KiB more of a workgroup's 64 KiB), at most 16 rows, and at least 512 workgroups
    // where the grid allows
    {
        const int nsp = std::min(ngr, SP_MAX_SPREAD);
        int pick = 0, smallest = 1;
        for (int R = std::min(16, K2); R >= 1; --R) {
            if (K2 % R || (long)(nsp * 8 + 32) * R * K3 + 16L * K3 > 60000) continue;
            smallest = R;
            if (!pick && (long)K1 * (K2 / R) >= 512) pick = R;
        }
        G.rows = pick ? pick : smallest;
    }
    const long ng = (long)K1 * K2 * K3;
    if (B.n_cap < n || !B.gi) {
        sp_free(B.gi);
        sp_free(B.th);
        sp_free(B.bin_site);
        B.n_cap = 0;
        const size_t nn = (size_t)std::max(n, 1);
        SPCHK(mbgpu_alloc(&B.gi, nn));
        SPCHK(mbgpu_alloc(&B.th, nn * 9 * SPME_MAX_ORDER));
        SPCHK(mbgpu_alloc(&B.bin_site, nn));
        B.n_cap = n;
    }
    if (!B.bin_off) SPCHK(mbgpu_alloc(&B.bin_off, (size_t)SPME_MAX_GRID + 1));
    if (B.ng_cap < ng || B.ngr_cap < ngr) {
        const long cap = std::max(B.ng_cap, ng);
        const int gcap = std::max(B.ngr_cap, ngr);
        sp_free(B.grid);
        sp_free(B.part);
        B.ng_cap = 0;
        B.ngr_cap = 0;
        SPCHK(mbgpu_alloc(&B.grid, (size_t)cap * gcap));
        SPCHK(mbgpu_alloc(&B.part, (size_t)((cap + 255) / 256) * gcap));
        B.ng_cap = cap;
        B.ngr_cap = gcap;
    }
    if (!B.out) SPCHK(mbgpu_alloc(&B.out, (size_t)SP_MAX_GRIDS));
    if (!B.flag) SPCHK(mbgpu_alloc(&B.flag, 2));
    if (B.cfg[0] != order || B.cfg[1] != K1 || B.cfg[2] != K2 || B.cfg[3] != K3) {
        // |b(m)|^2 = 1 / |sum_{k=0}^{p-2} M_p(k+1) exp(2 pi i m k / K)|^2 (bspcoe.f); M_p at
        // the integers by the recursion of bspgen.f:27-100 at w = 0
        std::vector<double> mp(order + 1, 0.0), nx(order + 1);
        mp[1] = 1.0;   // M_2(1) = 1 (M_2(0) = M_2(2) = 0)
        for (int k = 3; k <= order; ++k) {
            for (int j = 0; j <= k; ++j)
                nx[j] = ((double)j * (j < (int)mp.size() ? mp[j] : 0.0) + (double)(k - j) * (j > 0 ? mp[j - 1] : 0.0)) /
                        (double)(k - 1);
            for (int j = 0; j <= order; ++j) mp[j] = j <= k ? nx[j] : 0.0;
        }
        std::vector<double> bm(K1 + K2 + K3);
        std::vector<double2> tw;
        int off = 0;
        for (int a = 0; a < 3; ++a) {
            const int K = grid[a];
            for (int m = 0; m < K; ++m) {
                double re = 0.0, im = 0.0;
                for (int k = 0; k <= order - 2; ++k) {
                    const double ang = 2.0 * M_PI * (double)(((long)m * k) % K) / (double)K;
                    re += mp[k + 1] * std::cos(ang);
                    im += mp[k + 1] * std::sin(ang);
                }
                bm[off + m] = 1.0 / (re * re + im * im);
            }
            off += K;
            B.twoff[a] = (int)tw.size();
            for (int t = 0; t < K; ++t) {
                const double ang = 2.0 * M_PI * (double)t / (double)K;
                tw.push_back(make_double2(std::cos(ang), -std::sin(ang)));
            }
        }
        sp_free(B.bmod);
        sp_free(B.tw);
        B.cfg[0] = 0;
        SPCHK(mbgpu_alloc(&B.bmod, bm.size()));
        SPCHK(mbgpu_alloc(&B.tw, tw.size()));
        SPCHK(hipMemcpy(B.bmod, bm.data(), sizeof(double) * bm.size(), hipMemcpyHostToDevice));
        SPCHK(hipMemcpy(B.tw, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice));
        B.cfg[0] = order;
        B.cfg[1] = K1;
        B.cfg[2] = K2;
        B.cfg[3] = K3;
    }
    // transform lines: axis 2 contiguous lines, axes 1 and 0 strided with neighbouring
    // lines side by side
    B.axis[2] = SpmeAxis{K3, 1, K1 * K2, K3, 1, 0, 0L, 0, {}};
    B.axis[1] = SpmeAxis{K2, K3, K3, 1, K1, 0, (long)K2 * K3, 0, {}};
    B.axis[0] = SpmeAxis{K1, K2 * K3, K2 * K3, 1, 1, 0, 0L, 0, {}};
    for (int a = 0; a < 3; ++a) sp_axis_plan(B.axis[a]);
    return 0;
}
