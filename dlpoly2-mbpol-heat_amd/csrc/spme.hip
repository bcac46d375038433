// spme.hip -- smooth particle-mesh Ewald reciprocal space for point charges and point
// dipoles: energy, potential, field and force at every site.
//
// The reference's SPME (ewald_spme.f, spme_for.f, bspgen.f, bspcoe.f) spreads charges with
// cardinal B-splines of order p onto a K1 x K2 x K3 grid, transforms it, multiplies by the
// influence function B(m) C(m) and gathers potential derivatives back; this file does the
// same with the dipoles carried in the spread density (mu . grad of the spline product),
// the reciprocal part of the induced-dipole Ewald sum (ewald1p.f), to which it converges.
//
// Device steps (one stream, no host round trip):
//   k_sp_zero       clear the fixed-point grid
//   k_sp_spread     one thread per (site, t1, t2), p points along the third axis each;
//                   contributions added as 64-bit fixed point (2^-44 e), so the grid does
//                   not depend on the order the atomics land in
//   k_sp_fft x3     radix-2 Stockham passes in LDS along each axis (the first pass reads
//                   the fixed-point grid); L lines of one axis per workgroup of 128 threads
//   k_sp_influence  Q^(m) *= B(m) C(m); U = 1/2 sum B C |Q^|^2 as block partials
//   k_sp_fft x3     backward passes (unnormalised, conjugate twiddles)
//   k_sp_gather     one wavefront per site: phi, grad phi and the Hessian of phi from the
//                   p^3 grid points, reduced over the lanes in a fixed order
//   k_sp_usum       the energy partials in a fixed order
// Coulomb constant 1 (energy e^2/A); the caller scales.
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

// M_p(w + t) (and M_p', M_p'' when asked) for t = 0..p-1 by the recursion in the order
// (bspgen.f:27-100): M_n(x) = [x M_{n-1}(x) + (n - x) M_{n-1}(x - 1)] / (n - 1) from
// M_2(w) = w, M_2(w + 1) = 1 - w; M_n' = M_{n-1}(x) - M_{n-1}(x - 1).
template <int NDER>
__device__ void sp_splines(int p, double w, double* a, double* d, double* d2)
{
    double m[SPME_MAX_ORDER], m1[SPME_MAX_ORDER], m2[SPME_MAX_ORDER];
    for (int t = 0; t < SPME_MAX_ORDER; ++t) m[t] = m1[t] = m2[t] = 0.0;
    m[0] = w;
    m[1] = 1.0 - w;
    for (int n = 3; n <= p; ++n) {
        if (n == p - 1)
            for (int t = 0; t < SPME_MAX_ORDER; ++t) m2[t] = m[t];
        if (n == p)
            for (int t = 0; t < SPME_MAX_ORDER; ++t) m1[t] = m[t];
        double prev = 0.0;
        for (int t = 0; t < n; ++t) {
            const double x = w + t;
            const double cur = t < n - 1 ? m[t] : 0.0;
            const double nx = (x * cur + ((double)n - x) * prev) / (double)(n - 1);
            prev = cur;
            m[t] = nx;
        }
    }
    for (int t = 0; t < p; ++t) {
        a[t] = m[t];
        if (NDER >= 1) d[t] = (t < p - 1 ? m1[t] : 0.0) - (t > 0 ? m1[t - 1] : 0.0);
        if (NDER >= 2)
            d2[t] = (t < p - 2 ? m2[t] : 0.0) - 2.0 * (t > 0 ? m2[t - 1] : 0.0) + (t > 1 ? m2[t - 2] : 0.0);
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

__global__ __launch_bounds__(256) void k_sp_zero(unsigned long long* __restrict__ g, long ng)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < ng; i += (long)gridDim.x * 256) g[i] = 0ull;
}

// one block holds spb = 256 / p^2 sites; LDS: per (site, axis) the p splines and their
// derivatives, the grid index of t = 0 and the weight K_a (a*_a . mu)
__global__ __launch_bounds__(256) void k_sp_spread(SpmeGeom G, int n, const double* __restrict__ x,
                                                   const double* __restrict__ y, const double* __restrict__ z,
                                                   const double* __restrict__ q, const double* __restrict__ mu,
                                                   unsigned long long* __restrict__ grid, int* __restrict__ flag)
{
    __shared__ double th[256 / 16][3][2][SPME_MAX_ORDER];
    __shared__ double wv[256 / 16][3];
    __shared__ int gs[256 / 16][3];
    __shared__ int ok[256 / 16][3];
    const int p = G.p, pp = p * p, spb = 256 / pp;
    const int tid = threadIdx.x;
    const int s0 = blockIdx.x * spb;
    if (tid < 3 * spb) {
        const int s = tid / 3, a = tid % 3, i = s0 + s;
        long g0 = 0;
        double w = 0.0;
        const bool good = i < n && sp_frac(G, a, x[i], y[i], z[i], g0, w);
        if (good) {
            sp_splines<1>(p, w, th[s][a][0], th[s][a][1], nullptr);
            gs[s][a] = sp_wrap(g0, G.K[a]);
            wv[s][a] = mu ? G.K[a] * (G.rv[a][0] * mu[3 * i] + G.rv[a][1] * mu[3 * i + 1] + G.rv[a][2] * mu[3 * i + 2])
                          : 0.0;
        } else if (i < n) {
            flag[0] = 1;
        }
        ok[s][a] = good ? 1 : 0;
    }
    __syncthreads();
    if (tid >= spb * pp) return;
    const int s = tid / pp, t1 = (tid / p) % p, t2 = tid % p, i = s0 + s;
    if (i >= n || !(ok[s][0] & ok[s][1] & ok[s][2])) return;
    const double qi = q[i];
    const double a1 = th[s][0][0][t1], d1 = th[s][0][1][t1], a2 = th[s][1][0][t2], d2 = th[s][1][1][t2];
    const double c12 = qi * a1 * a2 + wv[s][0] * d1 * a2 + wv[s][1] * a1 * d2;
    const double w3 = wv[s][2] * a1 * a2;
    const int K1 = G.K[0], K2 = G.K[1], K3 = G.K[2];
    int g1 = gs[s][0] - t1, g2 = gs[s][1] - t2;
    g1 += g1 < 0 ? K1 : 0;
    g2 += g2 < 0 ? K2 : 0;
    const long row = ((long)g1 * K2 + g2) * K3;
    bool big = false;
    for (int t3 = 0; t3 < p; ++t3) {
        const double c = c12 * th[s][2][0][t3] + w3 * th[s][2][1][t3];
        big |= !(fabs(c) < 1024.0);
        int g3 = gs[s][2] - t3;
        g3 += g3 < 0 ? K3 : 0;
        atomicAdd(grid + row + g3, (unsigned long long)llrint(c * SP_FIX));
    }
    if (big) flag[0] = 2;
}

// Radix-2 Stockham transform of L lines of length K along one axis (K a power of two).
// Line l of block (o, ib) starts at o * so + (ib * L + l) * si, element e at + e * se.
// FIX: the input is the fixed-point grid (first forward pass).
template <bool FIX>
__global__ __launch_bounds__(128) void k_sp_fft(double2* __restrict__ g, const unsigned long long* __restrict__ gi,
                                                SpmeAxis A, const double2* __restrict__ tw, int inverse)
{
    __shared__ double2 buf[2][256];
    const int K = A.K, half = K >> 1, L = A.L;
    const int nib = A.ni / L;
    const int o = blockIdx.x / nib, ib = blockIdx.x - o * nib;
    const long base = (long)o * A.so + (long)ib * L * A.si;
    const int LK = L * K;
    for (int idx = threadIdx.x; idx < LK; idx += 128) {
        int l, e;
        if (A.se == 1) {
            l = idx / K;
            e = idx - l * K;
        } else {
            e = idx / L;
            l = idx - e * L;
        }
        const long at = base + (long)l * A.si + (long)e * A.se;
        double2 v;
        if (FIX) {
            v.x = (double)(long long)gi[at] * SP_UNFIX;
            v.y = 0.0;
        } else {
            v = g[at];
        }
        buf[0][l * K + e] = v;
    }
    __syncthreads();
    int src = 0;
    for (int Ns = 1; Ns < K; Ns <<= 1) {
        for (int b = threadIdx.x; b < L * half; b += 128) {
            const int l = b / half, j = b - l * half;
            const double2* in = buf[src] + l * K;
            double2* out = buf[src ^ 1] + l * K;
            const double2 v0 = in[j], v1 = in[j + half];
            const int k = j & (Ns - 1);
            double2 w = tw[k * (half / Ns)];
            if (inverse) w.y = -w.y;
            const double tr = v1.x * w.x - v1.y * w.y, ti = v1.x * w.y + v1.y * w.x;
            const int d = ((j - k) << 1) + k;
            out[d] = make_double2(v0.x + tr, v0.y + ti);
            out[d + Ns] = make_double2(v0.x - tr, v0.y - ti);
        }
        __syncthreads();
        src ^= 1;
    }
    for (int idx = threadIdx.x; idx < LK; idx += 128) {
        int l, e;
        if (A.se == 1) {
            l = idx / K;
            e = idx - l * K;
        } else {
            e = idx / L;
            l = idx - e * L;
        }
        g[base + (long)l * A.si + (long)e * A.se] = buf[src][l * K + e];
    }
}

// Q^(m) *= B(m) C(m), C(m) = exp(-pi^2 m^2 / alpha^2) / (pi V m^2), m = sum_a m_a a*_a with
// m_a folded into (-K_a/2, K_a/2]; block partials of 1/2 B C |Q^|^2
__global__ __launch_bounds__(256) void k_sp_influence(SpmeGeom G, double2* __restrict__ g,
                                                      const double* __restrict__ bmod, double* __restrict__ part)
{
    __shared__ double red[4];
    const int K1 = G.K[0], K2 = G.K[1], K3 = G.K[2];
    const long ng = (long)K1 * K2 * K3;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    if (i < ng) {
        const int m3 = (int)(i % K3), m2 = (int)((i / K3) % K2), m1 = (int)(i / ((long)K2 * K3));
        const int a1 = m1 <= K1 / 2 ? m1 : m1 - K1, a2 = m2 <= K2 / 2 ? m2 : m2 - K2, a3 = m3 <= K3 / 2 ? m3 : m3 - K3;
        double bc = 0.0;
        if (a1 || a2 || a3) {
            double mv[3];
            for (int d = 0; d < 3; ++d) mv[d] = a1 * G.rv[0][d] + a2 * G.rv[1][d] + a3 * G.rv[2][d];
            const double msq = mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2];
            bc = bmod[m1] * bmod[K1 + m2] * bmod[K1 + K2 + m3] * exp(-SP_PI * SP_PI * msq / (G.alpha * G.alpha)) /
                 (SP_PI * G.vol * msq);
        }
        double2 v = g[i];
        e = 0.5 * bc * (v.x * v.x + v.y * v.y);
        v.x *= bc;
        v.y *= bc;
        g[i] = v;
    }
    e = sp_wsum(e);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ __launch_bounds__(256) void k_sp_usum(int nb, const double* __restrict__ part, double* __restrict__ out)
{
    __shared__ double red[4];
    double a = 0.0;
    for (int k = threadIdx.x; k < nb; k += 256) a += part[k];
    a = sp_wsum(a);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = ((red[0] + red[1]) + red[2]) + red[3];
}

// one wavefront per site (4 per block): the lanes take the p^3 grid points in turn
__global__ __launch_bounds__(256) void k_sp_gather(SpmeGeom G, int n, const double* __restrict__ x,
                                                   const double* __restrict__ y, const double* __restrict__ z,
                                                   const double* __restrict__ q, const double* __restrict__ mu,
                                                   const double2* __restrict__ g, double* __restrict__ phi,
                                                   double* __restrict__ ef, double* __restrict__ f)
{
    __shared__ double th[4][3][3][SPME_MAX_ORDER];
    __shared__ int gs[4][3];
    __shared__ int ok[4][3];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + wv;
    const int p = G.p;
    if (lane < 3) {
        long g0 = 0;
        double w = 0.0;
        const bool good = i < n && sp_frac(G, lane, x[i], y[i], z[i], g0, w);
        if (good) {
            sp_splines<2>(p, w, th[wv][lane][0], th[wv][lane][1], th[wv][lane][2]);
            gs[wv][lane] = sp_wrap(g0, G.K[lane]);
        }
        ok[wv][lane] = good ? 1 : 0;
    }
    __syncthreads();
    if (i >= n || !(ok[wv][0] & ok[wv][1] & ok[wv][2])) return;
    const int K1 = G.K[0], K2 = G.K[1], K3 = G.K[2];
    double s[10];
    for (int k = 0; k < 10; ++k) s[k] = 0.0;
    const int np3 = p * p * p;
    for (int fidx = lane; fidx < np3; fidx += 64) {
        const int t1 = fidx / (p * p), t2 = (fidx / p) % p, t3 = fidx % p;
        int g1 = gs[wv][0] - t1, g2 = gs[wv][1] - t2, g3 = gs[wv][2] - t3;
        g1 += g1 < 0 ? K1 : 0;
        g2 += g2 < 0 ? K2 : 0;
        g3 += g3 < 0 ? K3 : 0;
        const double v = g[((long)g1 * K2 + g2) * K3 + g3].x;
        const double A1 = th[wv][0][0][t1], D1 = th[wv][0][1][t1], E1 = th[wv][0][2][t1];
        const double A2 = th[wv][1][0][t2], D2 = th[wv][1][1][t2], E2 = th[wv][1][2][t2];
        const double A3 = th[wv][2][0][t3], D3 = th[wv][2][1][t3], E3 = th[wv][2][2][t3];
        s[0] += v * A1 * A2 * A3;
        s[1] += v * D1 * A2 * A3;
        s[2] += v * A1 * D2 * A3;
        s[3] += v * A1 * A2 * D3;
        s[4] += v * E1 * A2 * A3;   // uu 00
        s[5] += v * A1 * E2 * A3;   // 11
        s[6] += v * A1 * A2 * E3;   // 22
        s[7] += v * D1 * D2 * A3;   // 01
        s[8] += v * D1 * A2 * D3;   // 02
        s[9] += v * A1 * D2 * D3;   // 12
    }
    for (int k = 0; k < 10; ++k) s[k] = sp_wsum(s[k]);
    if (lane != 0) return;
    // d/dr_d = sum_a K_a a*_a[d] d/du_a
    double ga[3][3];
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) ga[a][d] = G.K[a] * G.rv[a][d];
    const double hu[3][3] = {{s[4], s[7], s[8]}, {s[7], s[5], s[9]}, {s[8], s[9], s[6]}};
    double gr[3];
    for (int d = 0; d < 3; ++d) gr[d] = ga[0][d] * s[1] + ga[1][d] * s[2] + ga[2][d] * s[3];
    if (phi) phi[i] = s[0];
    if (ef)
        for (int d = 0; d < 3; ++d) ef[3 * i + d] = -gr[d];
    if (f) {
        const double qi = q[i];
        double m[3] = {0.0, 0.0, 0.0};
        if (mu)
            for (int d = 0; d < 3; ++d) m[d] = mu[3 * i + d];
        for (int d = 0; d < 3; ++d) {
            double hm = 0.0;
            for (int e = 0; e < 3; ++e) {
                double h = 0.0;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) h += ga[a][d] * ga[b][e] * hu[a][b];
                hm += h * m[e];
            }
            f[3 * i + d] = -qi * gr[d] - hm;
        }
    }
}

void sp_fft3(SpmeBuffers& B, const SpmeGeom& G, bool inverse, hipStream_t s)
{
    // axes in the order 2, 1, 0 forward (the first reads the fixed-point grid), 0, 1, 2 back
    for (int k = 0; k < 3; ++k) {
        const int a = inverse ? k : 2 - k;
        const SpmeAxis& A = B.axis[a];
        const unsigned nblk = (unsigned)((long)A.no * (A.ni / A.L));
        const double2* tw = B.tw + B.twoff[a];
        if (!inverse && k == 0)
            hipLaunchKernelGGL(k_sp_fft<true>, dim3(nblk), dim3(128), 0, s, B.grid, B.fixed, A, tw, 0);
        else
            hipLaunchKernelGGL(k_sp_fft<false>, dim3(nblk), dim3(128), 0, s, B.grid, B.fixed, A, tw, inverse ? 1 : 0);
    }
}

}  // namespace

int spme_launch(SpmeBuffers& B, const SpmeGeom& G, int n, const double* x, const double* y, const double* z,
                const double* q, const double* mu, double* phi, double* ef, double* f, hipStream_t s)
{
    const long ng = (long)G.K[0] * G.K[1] * G.K[2];
    const int zb = (int)std::min<long>(2048, (ng + 255) / 256);
    hipLaunchKernelGGL(k_sp_zero, dim3(zb), dim3(256), 0, s, B.fixed, ng);
    hipLaunchKernelGGL(k_sp_zero, dim3(1), dim3(256), 0, s, reinterpret_cast<unsigned long long*>(B.flag), 1L);
    const int spb = 256 / (G.p * G.p);
    if (n > 0)
        hipLaunchKernelGGL(k_sp_spread, dim3((n + spb - 1) / spb), dim3(256), 0, s, G, n, x, y, z, q, mu, B.fixed,
                           B.flag);
    sp_fft3(B, G, false, s);
    const int nb = (int)((ng + 255) / 256);
    hipLaunchKernelGGL(k_sp_influence, dim3(nb), dim3(256), 0, s, G, B.grid, B.bmod, B.part);
    hipLaunchKernelGGL(k_sp_usum, dim3(1), dim3(256), 0, s, nb, B.part, B.out);
    sp_fft3(B, G, true, s);
    if (n > 0 && (phi || ef || f))
        hipLaunchKernelGGL(k_sp_gather, dim3((n + 3) / 4), dim3(256), 0, s, G, n, x, y, z, q, mu, B.grid, phi, ef, f);
    return (int)hipGetLastError();
}
