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
//   k_sp_prep       per (site, axis): grid index, M_p, M_p', M_p'' and the dipole weight
//   k_sp_tile       one workgroup per (plane, 16 rows) of the grid: the tile in LDS as
//                   64-bit fixed point (2^-44 e), contributions of every site whose
//                   support meets it added with LDS atomics (integer adds, so the grid
//                   does not depend on their order), then written as complex values
//   k_sp_fft x3     radix-2 Stockham passes in LDS along each axis; L lines of one axis
//                   per workgroup of 128 threads (L K = 256 elements when the axis allows)
//   k_sp_influence  Q^(m) *= B(m) C(m); U = 1/2 sum B C |Q^|^2 as block partials
//   k_sp_fft x3     backward passes (unnormalised, conjugate twiddles)
//   k_sp_gather     one wavefront per site: phi, grad phi and (for forces) the Hessian of
//                   phi from the p^3 grid points, reduced over the lanes in a fixed order
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

// Per (site, axis): the grid index of t = 0 (wrapped), M_p, M_p', M_p'' at the p points and
// the dipole weight K_a (a*_a . mu); ok = 0 (and flag 1) for a position that cannot index
// the grid.  Layout: gi[i] = (g1, g2, g3, ok); th[i][a][k][t] (k: M, M', M''); cw[i] =
// (q, w1, w2, w3).
__global__ __launch_bounds__(256) void k_sp_prep(SpmeGeom G, int n, const double* __restrict__ x,
                                                 const double* __restrict__ y, const double* __restrict__ z,
                                                 const double* __restrict__ q, const double* __restrict__ mu,
                                                 int4* __restrict__ gi, double* __restrict__ th,
                                                 double4* __restrict__ cw, int* __restrict__ flag)
{
    __shared__ int okv[256];
    __shared__ int gv[256];
    __shared__ double wvv[256];
    const int t = blockIdx.x * 255 + threadIdx.x;   // (site, axis) pairs, 255 = 3 * 85 per block
    const int i = t / 3, a = t - 3 * (t / 3);
    const int p = G.p;
    const bool live = threadIdx.x < 255 && i < n;
    int good = 0, g = 0;
    double wa = 0.0;
    if (live) {
        long g0 = 0;
        double w = 0.0;
        if (sp_frac(G, a, x[i], y[i], z[i], g0, w)) {
            good = 1;
            g = sp_wrap(g0, G.K[a]);
            double* T = th + ((long)i * 3 + a) * 3 * SPME_MAX_ORDER;
            sp_splines<2>(p, w, T, T + SPME_MAX_ORDER, T + 2 * SPME_MAX_ORDER);
            if (mu) wa = G.K[a] * (G.rv[a][0] * mu[3 * i] + G.rv[a][1] * mu[3 * i + 1] + G.rv[a][2] * mu[3 * i + 2]);
        } else {
            flag[0] = 1;
        }
    }
    okv[threadIdx.x] = good;
    gv[threadIdx.x] = g;
    wvv[threadIdx.x] = wa;
    __syncthreads();
    if (live && a == 0) {
        const int b = threadIdx.x;
        gi[i] = make_int4(gv[b], gv[b + 1], gv[b + 2], okv[b] & okv[b + 1] & okv[b + 2]);
        cw[i] = make_double4(q[i], wvv[b], wvv[b + 1], wvv[b + 2]);
    }
}

// Spread by tiles: a workgroup owns plane g1 and rows r0..r0+R-1 of the grid (all g3) in
// LDS as 64-bit fixed point (2^-44 e), scans every site, lists those whose support meets
// the tile and adds their contributions with LDS atomics (integer adds: the tile does not
// depend on their order), then writes the tile as complex values.  Contribution at
// (t1, t2, t3): q M1 M2 M3 + w1 M1' M2 M3 + w2 M1 M2' M3 + w3 M1 M2 M3'.
constexpr int SP_ROWS = 16;
__global__ __launch_bounds__(256) void k_sp_tile(SpmeGeom G, int n, const int4* __restrict__ gi,
                                                 const double* __restrict__ th, const double4* __restrict__ cw,
                                                 double2* __restrict__ grid, int* __restrict__ flag)
{
    __shared__ unsigned long long tile[SP_ROWS * SPME_MAX_GRID];
    __shared__ int list[256], lt1[256];
    __shared__ int cnt;
    const int K1 = G.K[0], K2 = G.K[1], K3 = G.K[2], p = G.p;
    const int R = min(SP_ROWS, K2), nrb = K2 / R;
    const int g1 = blockIdx.x / nrb, r0 = (blockIdx.x - g1 * nrb) * R;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < R * K3; idx += 256) tile[idx] = 0ull;
    bool big = false;
    for (int base = 0; base < n; base += 256) {
        if (tid == 0) cnt = 0;
        __syncthreads();
        const int i = base + tid;
        if (i < n) {
            const int4 g = gi[i];
            int t1 = g.x - g1;
            t1 += t1 < 0 ? K1 : 0;
            if (g.w && t1 < p) {
                int d = g.y - r0;
                d += d < 0 ? K2 : 0;
                bool hit = false;
                for (int t2 = 0; t2 < p; ++t2) {
                    int r = d - t2;
                    r += r < 0 ? K2 : 0;
                    hit |= r < R;
                }
                if (hit) {
                    const int k = atomicAdd(&cnt, 1);
                    list[k] = i;
                    lt1[k] = t1;
                }
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
            const double4 c = cw[i2];
            const double a1 = T[t1], d1 = T[SPME_MAX_ORDER + t1];
            const double a2 = T[3 * SPME_MAX_ORDER + t2], d2 = T[4 * SPME_MAX_ORDER + t2];
            const double c12 = c.x * a1 * a2 + c.y * d1 * a2 + c.z * a1 * d2;
            const double w3 = c.w * a1 * a2;
            const double* T3 = T + 6 * SPME_MAX_ORDER;
            unsigned long long* row = tile + r * K3;
            for (int t3 = 0; t3 < p; ++t3) {
                const double v = c12 * T3[t3] + w3 * T3[SPME_MAX_ORDER + t3];
                big |= !(fabs(v) < 1024.0);
                int g3 = g.z - t3;
                g3 += g3 < 0 ? K3 : 0;
                atomicAdd(row + g3, (unsigned long long)llrint(v * SP_FIX));
            }
        }
        __syncthreads();
    }
    if (big) flag[0] = 2;
    for (int idx = tid; idx < R * K3; idx += 256) {
        const int r = idx / K3, c3 = idx - r * K3;
        grid[((long)g1 * K2 + r0 + r) * K3 + c3] = make_double2((double)(long long)tile[idx] * SP_UNFIX, 0.0);
    }
}

// Radix-2 Stockham transform of L lines of length K along one axis (K a power of two).
// Line l of block (o, ib) starts at o * so + (ib * L + l) * si, element e at + e * se.
__global__ __launch_bounds__(SP_FFT_THREADS) void k_sp_fft(double2* __restrict__ g, SpmeAxis A, const double2* __restrict__ tw,
                                                int inverse)
{
    __shared__ double2 buf[2][SP_FFT_ELEMS > SPME_MAX_GRID ? SP_FFT_ELEMS : SPME_MAX_GRID];
    const int K = A.K, half = K >> 1, L = A.L;
    const int nib = A.ni / L;
    const int o = blockIdx.x / nib, ib = blockIdx.x - o * nib;
    const long base = (long)o * A.so + (long)ib * L * A.si;
    const int LK = L * K;
    for (int idx = threadIdx.x; idx < LK; idx += SP_FFT_THREADS) {
        int l, e;
        if (A.se == 1) {
            l = idx / K;
            e = idx - l * K;
        } else {
            e = idx / L;
            l = idx - e * L;
        }
        buf[0][l * K + e] = g[base + (long)l * A.si + (long)e * A.se];
    }
    __syncthreads();
    int src = 0;
    for (int Ns = 1; Ns < K; Ns <<= 1) {
        for (int b = threadIdx.x; b < L * half; b += SP_FFT_THREADS) {
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
    for (int idx = threadIdx.x; idx < LK; idx += SP_FFT_THREADS) {
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

// one wavefront per site (4 per block): the lanes take the p^3 grid points in turn; HESS:
// the second derivatives too (forces asked for)
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
    const int p = G.p;
    const int K1 = G.K[0], K2 = G.K[1], K3 = G.K[2];
    const double* T = th + (long)i * 9 * SPME_MAX_ORDER;
    constexpr int MO = SPME_MAX_ORDER;
    double s[10];
    for (int k = 0; k < 10; ++k) s[k] = 0.0;
    const int np3 = p * p * p;
    for (int fidx = lane; fidx < np3; fidx += 64) {
        const int t1 = fidx / (p * p), t2 = (fidx / p) % p, t3 = fidx % p;
        int g1 = gg.x - t1, g2 = gg.y - t2, g3 = gg.z - t3;
        g1 += g1 < 0 ? K1 : 0;
        g2 += g2 < 0 ? K2 : 0;
        g3 += g3 < 0 ? K3 : 0;
        const double v = g[((long)g1 * K2 + g2) * K3 + g3].x;
        const double A1 = T[t1], D1 = T[MO + t1], E1 = T[2 * MO + t1];
        const double A2 = T[3 * MO + t2], D2 = T[4 * MO + t2], E2 = T[5 * MO + t2];
        const double A3 = T[6 * MO + t3], D3 = T[7 * MO + t3], E3 = T[8 * MO + t3];
        s[0] += v * A1 * A2 * A3;
        s[1] += v * D1 * A2 * A3;
        s[2] += v * A1 * D2 * A3;
        s[3] += v * A1 * A2 * D3;
        if (HESS) {
            s[4] += v * E1 * A2 * A3;   // second derivatives in u: 00
            s[5] += v * A1 * E2 * A3;   // 11
            s[6] += v * A1 * A2 * E3;   // 22
            s[7] += v * D1 * D2 * A3;   // 01
            s[8] += v * D1 * A2 * D3;   // 02
            s[9] += v * A1 * D2 * D3;   // 12
        }
    }
    for (int k = 0; k < (HESS ? 10 : 4); ++k) s[k] = sp_wsum(s[k]);
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
    if (HESS && f) {
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
    // axes in the order 2, 1, 0 forward, 0, 1, 2 back
    for (int k = 0; k < 3; ++k) {
        const int a = inverse ? k : 2 - k;
        const SpmeAxis& A = B.axis[a];
        const unsigned nblk = (unsigned)((long)A.no * (A.ni / A.L));
        hipLaunchKernelGGL(k_sp_fft, dim3(nblk), dim3(SP_FFT_THREADS), 0, s, B.grid, A, B.tw + B.twoff[a], inverse ? 1 : 0);
    }
}

}  // namespace

int spme_launch(SpmeBuffers& B, const SpmeGeom& G, int n, const double* x, const double* y, const double* z,
                const double* q, const double* mu, double* phi, double* ef, double* f, hipStream_t s)
{
    const long ng = (long)G.K[0] * G.K[1] * G.K[2];
    hipLaunchKernelGGL(k_sp_zero, dim3(1), dim3(256), 0, s, reinterpret_cast<unsigned long long*>(B.flag), 1L);
    if (n > 0)
        hipLaunchKernelGGL(k_sp_prep, dim3((unsigned)((3L * n + 254) / 255)), dim3(256), 0, s, G, n, x, y, z, q, mu,
                           B.gi, B.th, B.cw, B.flag);
    const int R = std::min(SP_ROWS, G.K[1]);
    hipLaunchKernelGGL(k_sp_tile, dim3((unsigned)(G.K[0] * (G.K[1] / R))), dim3(256), 0, s, G, n, B.gi, B.th, B.cw,
                       B.grid, B.flag);
    sp_fft3(B, G, false, s);
    const int nb = (int)((ng + 255) / 256);
    hipLaunchKernelGGL(k_sp_influence, dim3(nb), dim3(256), 0, s, G, B.grid, B.bmod, B.part);
    hipLaunchKernelGGL(k_sp_usum, dim3(1), dim3(256), 0, s, nb, B.part, B.out);
    sp_fft3(B, G, true, s);
    if (n > 0 && f)
        hipLaunchKernelGGL(k_sp_gather<true>, dim3((n + 3) / 4), dim3(256), 0, s, G, n, B.gi, B.th, q, mu, B.grid, phi,
                           ef, f);
    else if (n > 0 && (phi || ef))
        hipLaunchKernelGGL(k_sp_gather<false>, dim3((n + 3) / 4), dim3(256), 0, s, G, n, B.gi, B.th, q, mu, B.grid, phi,
                           ef, f);
    return (int)hipGetLastError();
}
