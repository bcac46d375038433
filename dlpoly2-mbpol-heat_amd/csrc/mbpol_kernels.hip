// mbpol_kernels.hip -- MB-pol 2B/3B force path on gfx950: monomer imaging, dimer and
// trimer enumeration, cluster set-up (extra points and polynomial variables), chain rule
// and switching, per-atom accumulation of forces / energies / heat-flux tensors.
//
// Reference semantics restated (file:line in the reference tree):
//   image_monomers   mbpol.f90:1275-1301     k_image_monomers
//   pair predicate   mbpol.f90:496-503       k_neighbours (|img(O_i-O_j)|^2 <= r2f^2)
//   triple predicate mbpol.f90:843-869       k_trimers (some monomer within r3f of both)
//   re-imaging       mbpol.f90:505-509,875-882  k_gather2 / k_gather3
//   2B leaf          mbpol-2b-poly.cpp:358-682  k_setup2 + pip2b + k_finish2
//   3B leaf          mbpol-3b-poly.cpp:148-312  k_setup3 + pip3b + k_finish3
//   accumulation     mbpol.f90:519-653, 896-1067  k_accum2 / k_accum3
//   heat-current bookkeeping  heatcurrent.f90 update_energy_mbpol / update_forces, contracted
//                    per atom as W_i = sum_j F_ij (x) (x_i - x_j) (heatcurrent.f90:716-719)
#include <stdio.h>

#include "mbgpu_internal.h"

namespace {

// ---------------------------------------------------------------- model constants ----
// 2B (mbpol-2b-poly.cpp:1848-1863)
constexpr double c2_k[8] = {
    -6.480884773303821e-01,  // 0 HH intra
    1.674518993682975e+00,   // 1 OH intra
    1.148231864355956e+00,   // 2 HH coul
    1.205989761123099e+00,   // 3 OH coul
    1.395357065790959e+00,   // 4 OO coul
    7.347036852042255e-01,   // 5 XH main
    7.998249864422826e-01,   // 6 XO main
    7.960663960630585e-01};  // 7 XX main
#define IN_PLANE_G (-9.721486914088159e-02)
#define OUT_PLANE_G (9.859272078406150e-02)
#define R2I 4.5

// the 31 variables (mbpol-2b-poly.cpp:413-445): kind (0 exp, 1 exp/r), k index, d0, sites
// sites: 0 Oa 1 Ha1 2 Ha2 3 Ob 4 Hb1 5 Hb2 6 Xa1 7 Xa2 8 Xb1 9 Xb2
struct V2Def { int kind, k, d0, s1, s2; };
constexpr V2Def c2_vars[31] = {
    {0, 0, 1, 1, 2}, {0, 0, 1, 4, 5}, {0, 1, 1, 0, 1}, {0, 1, 1, 0, 2}, {0, 1, 1, 3, 4},
    {0, 1, 1, 3, 5}, {1, 2, 4, 1, 4}, {1, 2, 4, 1, 5}, {1, 2, 4, 2, 4}, {1, 2, 4, 2, 5},
    {1, 3, 4, 0, 4}, {1, 3, 4, 0, 5}, {1, 3, 4, 3, 1}, {1, 3, 4, 3, 2}, {1, 4, 4, 0, 3},
    {0, 5, 4, 6, 4}, {0, 5, 4, 6, 5}, {0, 5, 4, 7, 4}, {0, 5, 4, 7, 5}, {0, 5, 4, 8, 1},
    {0, 5, 4, 8, 2}, {0, 5, 4, 9, 1}, {0, 5, 4, 9, 2}, {0, 6, 4, 0, 8}, {0, 6, 4, 0, 9},
    {0, 6, 4, 3, 6}, {0, 6, 4, 3, 7}, {0, 7, 4, 6, 8}, {0, 7, 4, 6, 9}, {0, 7, 4, 7, 8},
    {0, 7, 4, 7, 9}};

// 3B (mbpol-3b-poly.cpp:1488-1504): k and r0 per variable class
// class 0 HH intra, 1 OH intra, 2 HH, 3 OH, 4 OO
constexpr double c3_k[5] = {-2.254303257872797e+00, -2.564771901404151e-01, 4.247920544074333e-01,
                               8.128985941165371e-01, 3.749457984616480e-02};
constexpr double c3_r0[5] = {1.690594510379166e+00, -2.999999868517452e+00, 3.499031358429095e+00,
                                4.854042963514281e+00, 4.816312044947604e-08};
// the 36 variables (mbpol-3b-poly.cpp:170-206): class, atom 1, atom 2
// atoms: 0 Oa 1 Ha1 2 Ha2 3 Ob 4 Hb1 5 Hb2 6 Oc 7 Hc1 8 Hc2
struct V3Def { int cls, a1, a2; };
constexpr V3Def c3_vars[36] = {
    {0, 1, 2}, {0, 4, 5}, {0, 7, 8}, {1, 0, 1}, {1, 0, 2}, {1, 3, 4}, {1, 3, 5}, {1, 6, 7},
    {1, 6, 8}, {2, 1, 4}, {2, 1, 5}, {2, 1, 7}, {2, 1, 8}, {2, 2, 4}, {2, 2, 5}, {2, 2, 7},
    {2, 2, 8}, {2, 4, 7}, {2, 4, 8}, {2, 5, 7}, {2, 5, 8}, {3, 0, 4}, {3, 0, 5}, {3, 0, 7},
    {3, 0, 8}, {3, 3, 1}, {3, 3, 2}, {3, 3, 7}, {3, 3, 8}, {3, 6, 1}, {3, 6, 2}, {3, 6, 4},
    {3, 6, 5}, {4, 0, 3}, {4, 0, 6}, {4, 3, 6}};

// ------------------------------------------------------------------ small helpers ----
__device__ __forceinline__ void load_mon(const double* __restrict__ mon, int w, double* m)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = mon[9 * w + k];
}

// monomer::setup (mbpol-2b-poly.cpp:118-140)
__device__ __forceinline__ void extra_points(const double* ohh, double* oh1, double* oh2, double* x1,
                                             double* x2)
{
    for (int i = 0; i < 3; ++i) {
        oh1[i] = ohh[i + 3] - ohh[i];
        oh2[i] = ohh[i + 6] - ohh[i];
    }
    const double v[3] = {oh1[1] * oh2[2] - oh1[2] * oh2[1], oh1[2] * oh2[0] - oh1[0] * oh2[2],
                         oh1[0] * oh2[1] - oh1[1] * oh2[0]};
    for (int i = 0; i < 3; ++i) {
        const double in_plane = ohh[i] + 0.5 * IN_PLANE_G * (oh1[i] + oh2[i]);
        const double out_of_plane = OUT_PLANE_G * v[i];
        x1[i] = in_plane + out_of_plane;
        x2[i] = in_plane - out_of_plane;
    }
}

// monomer::grads (mbpol-2b-poly.cpp:144-176)
__device__ __forceinline__ void extra_point_grads(const double* oh1, const double* oh2, const double* g1,
                                                  const double* g2, double* grd)
{
    const double gm[3] = {g1[0] - g2[0], g1[1] - g2[1], g1[2] - g2[2]};
    const double t1[3] = {oh2[1] * gm[2] - oh2[2] * gm[1], oh2[2] * gm[0] - oh2[0] * gm[2],
                          oh2[0] * gm[1] - oh2[1] * gm[0]};
    const double t2[3] = {oh1[1] * gm[2] - oh1[2] * gm[1], oh1[2] * gm[0] - oh1[0] * gm[2],
                          oh1[0] * gm[1] - oh1[1] * gm[0]};
    for (int i = 0; i < 3; ++i) {
        const double gsum = g1[i] + g2[i];
        const double in_plane = 0.5 * IN_PLANE_G * gsum;
        const double gh1 = in_plane + OUT_PLANE_G * t1[i];
        const double gh2 = in_plane - OUT_PLANE_G * t2[i];
        grd[i + 0] += gsum - (gh1 + gh2);
        grd[i + 3] += gh1;
        grd[i + 6] += gh2;
    }
}

// f_switch (mbpol-2b-poly.cpp:340-354, mbpol-3b-poly.cpp:130-144)
__device__ __forceinline__ double f_switch(double r, double ri, double rf, double& g)
{
    if (r > rf) {
        g = 0.0;
        return 0.0;
    } else if (r > ri) {
        const double t1 = M_PI / (rf - ri);
        const double x = (r - ri) * t1;
        g = -sin(x) * t1 / 2.0;
        return (1.0 + cos(x)) / 2.0;
    }
    g = 0.0;
    return 1.0;
}

// 10 sites (30 coordinates) of a dimer: O,H,H of both waters then the 4 extra points
__device__ __forceinline__ void dimer_sites(const double* w, double* s, double* oha1, double* oha2,
                                           double* ohb1, double* ohb2)
{
    for (int k = 0; k < 18; ++k) s[k] = w[k];
    extra_points(s + 0, oha1, oha2, s + 18, s + 21);
    extra_points(s + 9, ohb1, ohb2, s + 24, s + 27);
}

// Block sum of n <= 16 per-thread values in a fixed order (wavefront shuffles, then the
// wavefronts in index order); thread q < n writes sum q to out[q].  Every thread of the
// block must call it.
__device__ __forceinline__ void block_reduce_store(const double* vals, int n, double* out)
{
    __shared__ double red[16][16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int q = 0; q < n; ++q) {
        double v = vals[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[wid][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < n) {
        double s = 0.0;
        const int nwv = (blockDim.x + 63) >> 6;
        for (int w = 0; w < nwv; ++w) s += red[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

}  // namespace

// ===================================================================== kernels ======

__global__ void k_zero_ll(long long* __restrict__ p, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void k_zero_i(int* __restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// image_monomers (mbpol.f90:1275-1301) on the waters gathered through listttm2.  Every
// water's entries are checked (water_base) before an atom array is read with them: a bad
// entry sets bit 1 of *err and the water reads atoms 0..2 instead, so that the evaluation
// stays in bounds until the host sees the flag and stops with an error.  The checked
// 0-based offsets go to wbase[nw] for the later kernels.
__global__ void k_image_monomers(int nw, int natms, const int* __restrict__ list, const double* __restrict__ x,
                                 const double* __restrict__ y, const double* __restrict__ z, CellGeom g,
                                 double* __restrict__ mon, int* __restrict__ wbase, int* __restrict__ err)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    int o = water_base(list, w, natms);
    if (o < 0) {
        atomicOr(err, 2);
        o = 0;
    }
    wbase[w] = o;
    double m[9];
    for (int a = 0; a < 3; ++a) {
        m[3 * a] = x[o + a];
        m[3 * a + 1] = y[o + a];
        m[3 * a + 2] = z[o + a];
    }
    for (int h = 1; h < 3; ++h) {
        double dx = m[3 * h] - m[0], dy = m[3 * h + 1] - m[1], dz = m[3 * h + 2] - m[2];
        image_vec(g, dx, dy, dz);
        m[3 * h] = m[0] + dx;
        m[3 * h + 1] = m[1] + dy;
        m[3 * h + 2] = m[2] + dz;
    }
    for (int k = 0; k < 9; ++k) mon[9 * w + k] = m[k];
}

// Ascending order of the first len (<= 4 * 64) entries of row (distinct values), in place,
// by one wavefront: each lane's entries get their rank among all entries.  Every lane reads
// the whole row before any lane writes it.
__device__ void wave_sort_row(int* __restrict__ row, int len)
{
    const int lane = threadIdx.x & 63;
    __threadfence_block();   // the row's entries were stored by other lanes of this wavefront
    int v[4], r[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = lane + 64 * u;
        v[u] = e < len ? row[e] : 0;
        r[u] = 0;
    }
    for (int k = 0; k < len; ++k) {
        const int w = row[k];
#pragma unroll
        for (int u = 0; u < 4; ++u) r[u] += w < v[u];
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (lane + 64 * u < len) row[r[u]] = v[u];
}

// O-O neighbour lists.  One wavefront per water i scans its candidate partners in chunks of
// 64 -- every water, or, when the box holds three link cells per axis (cells.on), the
// waters of the 27 cells around i (the link-cell Verlet build of parlink.f) -- and a
// ballot + popcount compacts the hits; the rows are in ascending partner order (sorted after
// a link-cell scan).
//   list 2:  j > i with |img(O_i - O_j)|^2 <= r2f^2          (do_dimer predicate)
//   list 2l: j < i with the same predicate (the partners of which i is the second water)
//   list 3:  j != i with sqrt(|img(O_lo - O_hi)|^2) <= r3f    (do_trimer edge test)
__global__ __launch_bounds__(64) void k_neighbours(int nw, const double* __restrict__ mon, CellGeom g, CellView cv,
                                                   int cap2, int* __restrict__ nbr2, int* __restrict__ cnt2,
                                                   int* __restrict__ nbr2l, int* __restrict__ cnt2l,
                                                   int cap3, int* __restrict__ nbr3, int* __restrict__ cnt3,
                                                   int* __restrict__ err)
{
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    if (i >= nw) return;
    int n2 = 0, n2l = 0, n3 = 0;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    auto take = [&](int j) {
        bool in2 = false, in2l = false, in3 = false;
        if (j >= 0 && j != i) {
            const int lo = min(i, j), hi = max(i, j);
            double dx = mon[9 * lo] - mon[9 * hi];
            double dy = mon[9 * lo + 1] - mon[9 * hi + 1];
            double dz = mon[9 * lo + 2] - mon[9 * hi + 2];
            image_vec(g, dx, dy, dz);
            const double r2 = norm2_nc(dx, dy, dz);
            const bool pair = !(r2 > MBGPU_R2F * MBGPU_R2F);
            in2 = (j > i) && pair;
            in2l = (j < i) && pair;
            in3 = !(sqrt(r2) > MBGPU_R3F);
        }
        const unsigned long long b2 = __ballot(in2), b2l = __ballot(in2l), b3 = __ballot(in3);
        if (in2) {
            const int p = n2 + __popcll(b2 & below);
            if (p < cap2) nbr2[(long)i * cap2 + p] = j;
        }
        if (in2l) {
            const int p = n2l + __popcll(b2l & below);
            if (p < cap2) nbr2l[(long)i * cap2 + p] = j;
        }
        if (in3) {
            const int p = n3 + __popcll(b3 & below);
            if (p < cap3) nbr3[(long)i * cap3 + p] = j;
        }
        n2 += __popcll(b2);
        n2l += __popcll(b2l);
        n3 += __popcll(b3);
    };
    if (cv.on) {
        for_neighbour_cells(cv, i, [&](int c0, int c1) {
            for (int base = c0; base < c1; base += 64) take(base + lane < c1 ? cv.idx[base + lane] : -1);
        });
        wave_sort_row(nbr2 + (long)i * cap2, min(n2, cap2));
        wave_sort_row(nbr2l + (long)i * cap2, min(n2l, cap2));
        wave_sort_row(nbr3 + (long)i * cap3, min(n3, cap3));
    } else {
        for (int base = 0; base < nw; base += 64) take(base + lane < nw ? base + lane : -1);
    }
    if (lane == 0) {
        cnt2[i] = n2;
        cnt2l[i] = n2l;
        cnt3[i] = n3;
        if (n2 > cap2 || n2l > cap2 || n3 > cap3) atomicOr(err, 1);
    }
}

// Trimer enumeration around a centre water c: every pair a < b of c's short neighbours
// forms a triple with two short edges (c-a, c-b).  A triple whose third edge is short too
// is reached from all three vertices and kept only from its smallest index.  This yields
// exactly the triples do_trimer keeps: at least two of the three O-O edges <= r3f.
// One wavefront per centre (at most 64 short neighbours, checked when the lists are
// built): row p of the pair table runs sequentially, the partner index q across the lanes.
// The count pass records, per row, the bit mask of kept partners and the number of kept
// pairs before the row, so that the position of any kept pair in c's output range is
// toff[c] + rowoff[c][p] + popcount(mask[c][p] below bit q); k_gather_water uses this to
// find the trimers of a water without atomics.
template <bool FILL>
__global__ __launch_bounds__(64) void k_trimers(int nw, const double* __restrict__ mon, CellGeom g,
                                                const int* __restrict__ nbr3, const int* __restrict__ cnt3,
                                                int* __restrict__ tcnt, const int* __restrict__ toff,
                                                unsigned long long* __restrict__ kmask, int* __restrict__ rowoff,
                                                int* __restrict__ trimers, long tcap)
{
    const int c = blockIdx.x;
    const int lane = threadIdx.x;
    if (c >= nw) return;
    const int n = min(cnt3[c], 64);
    const int* S = nbr3 + (long)c * 64;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int cnt = 0;
    const long o = FILL ? toff[c] : 0;
    for (int p = 0; p + 1 < n; ++p) {
        const int a = S[p];
        const int q = p + 1 + lane;
        bool keep = false;
        int b = 0;
        if (q < n) {
            b = S[q];
            double dx = mon[9 * a] - mon[9 * b];
            double dy = mon[9 * a + 1] - mon[9 * b + 1];
            double dz = mon[9 * a + 2] - mon[9 * b + 2];
            image_vec(g, dx, dy, dz);
            const double rab = sqrt(norm2_nc(dx, dy, dz));
            keep = (rab > MBGPU_R3F) || (c < a);
        }
        const unsigned long long bal = __ballot(keep);
        if (FILL) {
            if (keep) {
                int i0, i1, i2;
                if (c < a) { i0 = c; i1 = a; i2 = b; }
                else if (c < b) { i0 = a; i1 = c; i2 = b; }
                else { i0 = a; i1 = b; i2 = c; }
                const long t = o + cnt + __popcll(bal & below);
                if (t < tcap) {
                    trimers[3 * t] = i0;
                    trimers[3 * t + 1] = i1;
                    trimers[3 * t + 2] = i2;
                }
            }
        } else if (lane == 0) {
            kmask[(long)c * 64 + p] = bal << (p + 1);
            rowoff[(long)c * 64 + p] = cnt;
        }
        cnt += __popcll(bal);
    }
    if (!FILL && lane == 0) {
        tcnt[c] = cnt;
        for (int p = max(n - 1, 0); p < 64; ++p) {
            kmask[(long)c * 64 + p] = 0ull;
            rowoff[(long)c * 64 + p] = cnt;
        }
    }
}

// Exclusive prefix sum of n counts (one block of 1024 threads); total to *total.
__global__ __launch_bounds__(1024) void k_scan(const int* __restrict__ cnt, int n, int* __restrict__ off,
                                               int* __restrict__ total)
{
    __shared__ long wsum[16];
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    const int per = (n + 1023) / 1024;
    const int b = t * per, e = min(n, b + per);
    long s = 0;
    for (int i = b; i < e; ++i) s += cnt[i];
    // inclusive scan of s across the wavefront
    long inc = s;
    for (int d = 1; d < 64; d <<= 1) {
        const long v = __shfl_up(inc, d, 64);
        if (lane >= d) inc += v;
    }
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    if (t == 0) {
        long acc = 0;
        for (int w = 0; w < 16; ++w) {
            const long v = wsum[w];
            wsum[w] = acc;
            acc += v;
        }
        *total = (int)acc;
    }
    __syncthreads();
    long acc = wsum[wid] + inc - s;
    for (int i = b; i < e; ++i) {
        off[i] = (int)acc;
        acc += cnt[i];
    }
}

__global__ void k_fill_dimers(int nw, int cap2, const int* __restrict__ nbr2, const int* __restrict__ cnt2,
                              const int* __restrict__ off2, int* __restrict__ dimers, long dcap)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nw) return;
    const int n = min(cnt2[i], cap2);
    for (int q = 0; q < n; ++q) {
        const long d = (long)off2[i] + q;
        if (d < dcap) {
            dimers[2 * d] = i;
            dimers[2 * d + 1] = nbr2[(long)i * cap2 + q];
        }
    }
}

// Re-image a dimer around its second water (mbpol.f90:496-509); SoA [18][nd].
__global__ void k_gather2(const int* __restrict__ ncnt, int ncap, const int* __restrict__ dimers, const double* __restrict__ mon, CellGeom g,
                          double* __restrict__ clus)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const int nd = min(*ncnt, ncap);
    if (d >= nd) return;
    double w1[9], w2[9];
    load_mon(mon, dimers[2 * d], w1);
    load_mon(mon, dimers[2 * d + 1], w2);
    double r[3] = {w1[0] - w2[0], w1[1] - w2[1], w1[2] - w2[2]};
    image_vec(g, r[0], r[1], r[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        w1[3 + k] = (w1[3 + k] - w1[k]) + (w2[k] + r[k]);
        w1[6 + k] = (w1[6 + k] - w1[k]) + (w2[k] + r[k]);
        w1[k] = w2[k] + r[k];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        clus[(long)k * ncap + d] = w1[k];
        clus[(long)(9 + k) * ncap + d] = w2[k];
    }
}

// Re-image a trimer around its middle water (mbpol.f90:843-882); SoA [27][nt].
__global__ void k_gather3(const int* __restrict__ ncnt, int ncap, const int* __restrict__ trimers, const double* __restrict__ mon, CellGeom g,
                          double* __restrict__ clus)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = min(*ncnt, ncap);
    if (t >= nt) return;
    double w1[9], w2[9], w3[9];
    load_mon(mon, trimers[3 * t], w1);
    load_mon(mon, trimers[3 * t + 1], w2);
    load_mon(mon, trimers[3 * t + 2], w3);
    double ra[3] = {w1[0] - w2[0], w1[1] - w2[1], w1[2] - w2[2]};
    double rb[3] = {w2[0] - w3[0], w2[1] - w3[1], w2[2] - w3[2]};
    image_vec(g, ra[0], ra[1], ra[2]);
    image_vec(g, rb[0], rb[1], rb[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        w1[3 + k] = (w1[3 + k] - w1[k]) + (w2[k] + ra[k]);
        w1[6 + k] = (w1[6 + k] - w1[k]) + (w2[k] + ra[k]);
        w1[k] = w2[k] + ra[k];
        w3[3 + k] = (w3[3 + k] - w3[k]) + (w2[k] - rb[k]);
        w3[6 + k] = (w3[6 + k] - w3[k]) + (w2[k] - rb[k]);
        w3[k] = w2[k] - rb[k];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        clus[(long)k * ncap + t] = w1[k];
        clus[(long)(9 + k) * ncap + t] = w2[k];
        clus[(long)(18 + k) * ncap + t] = w3[k];
    }
}

// 2B variables (mbpol-2b-poly.cpp:358-445): SoA [31][nd]; xvf (may be NULL) receives the
// same values rounded to FP32 for the mixed-precision polynomial
__global__ void k_setup2(const int* __restrict__ ncnt, int ncap, const double* __restrict__ clus, double* __restrict__ xv, float* __restrict__ xvf)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const int nd = min(*ncnt, ncap);
    if (d >= nd) return;
    double w[18], s[30], a1[3], a2[3], b1[3], b2[3];
#pragma unroll
    for (int k = 0; k < 18; ++k) w[k] = clus[(long)k * ncap + d];
    dimer_sites(w, s, a1, a2, b1, b2);
#pragma unroll
    for (int n = 0; n < 31; ++n) {
        const V2Def v = c2_vars[n];
        const double dx = s[3 * v.s1] - s[3 * v.s2], dy = s[3 * v.s1 + 1] - s[3 * v.s2 + 1],
                     dz = s[3 * v.s1 + 2] - s[3 * v.s2 + 2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const double e = exp(c2_k[v.k] * ((double)v.d0 - r));
        const double val = v.kind ? e * (1.0 / r) : e;
        xv[(long)n * ncap + d] = val;
        if (xvf) xvf[(long)n * ncap + d] = (float)val;
    }
}

// Per-cluster virial and stress (do_dimer mbpol.f90:582-653, do_trimer :972-1067) with g
// in kcal/mol/A: vir = sum w.g*engunit, strs = -sum w (x) g*engunit (upper triangle).
template <int NC>
__device__ __forceinline__ void cluster_virial(const double* w, const double* g, double* red)
{
#pragma unroll
    for (int a = 0; a < NC / 3; ++a) {
        const double g0 = g[3 * a] * MBGPU_ENGUNIT, g1 = g[3 * a + 1] * MBGPU_ENGUNIT,
                     g2 = g[3 * a + 2] * MBGPU_ENGUNIT;
        red[1] += w[3 * a] * g0 + w[3 * a + 1] * g1 + w[3 * a + 2] * g2;
        red[2] -= w[3 * a] * g0;
        red[3] -= w[3 * a] * g1;
        red[4] -= w[3 * a] * g2;
        red[5] -= w[3 * a + 1] * g1;
        red[6] -= w[3 * a + 1] * g2;
        red[7] -= w[3 * a + 2] * g2;
    }
}

// Sum the slice partials, apply the chain rule through the variables and extra points and
// the switch (mbpol-2b-poly.cpp:447-682).  Record: E (kcal/mol) and dE/dx of the 6 atoms;
// block partials of energy, virial and stress.
// sum of the raw (un-imaged) positions of the atoms of the NMON waters ids[0..NMON)
template <int NMON>
__device__ __forceinline__ void cluster_raw_sum(const int* ids, const int* __restrict__ wbase,
                                                const double* __restrict__ x, const double* __restrict__ y,
                                                const double* __restrict__ z, double* S)
{
    S[0] = S[1] = S[2] = 0.0;
#pragma unroll
    for (int m = 0; m < NMON; ++m) {
        const int o = wbase[ids[m]];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            S[0] += x[o + a];
            S[1] += y[o + a];
            S[2] += z[o + a];
        }
    }
}

__global__ __launch_bounds__(128) void k_finish2(const int* __restrict__ ncnt, int ncap, int nsl, const double* __restrict__ clus,
                                                 const double* __restrict__ xv, const double* __restrict__ part,
                                                 double* __restrict__ erec, double* __restrict__ grec,
                                                 double* __restrict__ blk, const int* __restrict__ dimers,
                                                 const int* __restrict__ wbase, const double* __restrict__ x,
                                                 const double* __restrict__ y, const double* __restrict__ z,
                                                 double* __restrict__ srec)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const int nd = min(*ncnt, ncap);
    double red[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (d < nd && dimers) {
        double S[3];
        cluster_raw_sum<2>(dimers + 2L * d, wbase, x, y, z, S);
        srec[d] = S[0];
        srec[(long)ncap + d] = S[1];
        srec[2L * ncap + d] = S[2];
    }
    if (d < nd) {
        double w[18], s[30], a1[3], a2[3], b1[3], b2[3];
#pragma unroll
        for (int k = 0; k < 18; ++k) w[k] = clus[(long)k * ncap + d];
        const double dOO[3] = {w[0] - w[9], w[1] - w[10], w[2] - w[11]};
        const double rOO = sqrt(dOO[0] * dOO[0] + dOO[1] * dOO[1] + dOO[2] * dOO[2]);
        double gout[18];
        double Eout = 0.0;
        if (rOO > MBGPU_R2F) {  // the leaf returns 0 without touching the gradients
#pragma unroll
            for (int k = 0; k < 18; ++k) gout[k] = 0.0;
        } else {
            dimer_sites(w, s, a1, a2, b1, b2);
            double E = 0.0;
            for (int q = 0; q < nsl; ++q) E += part[((long)q * 32) * ncap + d];
            double xg[30];
#pragma unroll
            for (int k = 0; k < 30; ++k) xg[k] = 0.0;
#pragma unroll
            for (int n = 0; n < 31; ++n) {
                double gn = 0.0;
                for (int q = 0; q < nsl; ++q) gn += part[((long)q * 32 + 1 + n) * ncap + d];
                const V2Def v = c2_vars[n];
                const double dd[3] = {s[3 * v.s1] - s[3 * v.s2], s[3 * v.s1 + 1] - s[3 * v.s2 + 1],
                                      s[3 * v.s1 + 2] - s[3 * v.s2 + 2]};
                const double r = sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
                const double val = xv[(long)n * ncap + d];
                const double k = c2_k[v.k];
                double gg;
                if (v.kind) {
                    const double rinv = 1.0 / r;
                    gg = -(k + rinv) * val * rinv;
                } else {
                    gg = -k * val / r;
                }
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const double t = gn * (dd[i] * gg);
                    xg[3 * v.s1 + i] += t;
                    xg[3 * v.s2 + i] -= t;
                }
            }
            extra_point_grads(a1, a2, xg + 18, xg + 21, xg + 0);
            extra_point_grads(b1, b2, xg + 24, xg + 27, xg + 9);
            double gsw;
            const double sw = f_switch(rOO, R2I, MBGPU_R2F, gsw);
#pragma unroll
            for (int i = 0; i < 18; ++i) gout[i] = sw * xg[i];
            gsw *= E / rOO;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double t = gsw * dOO[i];
                gout[i] += t;
                gout[9 + i] -= t;
            }
            Eout = sw * E;
        }
        erec[d] = Eout;
#pragma unroll
        for (int k = 0; k < 18; ++k) grec[(long)k * ncap + d] = gout[k];
        red[0] = Eout * MBGPU_ENGUNIT;
        cluster_virial<18>(w, gout, red);
    }
    block_reduce_store(red, 8, blk + 8 * blockIdx.x);
}

// 3B variables (mbpol-3b-poly.cpp:170-206): SoA [36][nt]; xvf as in k_setup2
__global__ void k_setup3(const int* __restrict__ ncnt, int ncap, const double* __restrict__ clus, double* __restrict__ xv, float* __restrict__ xvf)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = min(*ncnt, ncap);
    if (t >= nt) return;
    double w[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) w[k] = clus[(long)k * ncap + t];
#pragma unroll
    for (int n = 0; n < 36; ++n) {
        const V3Def v = c3_vars[n];
        const double dx = w[3 * v.a1] - w[3 * v.a2], dy = w[3 * v.a1 + 1] - w[3 * v.a2 + 1],
                     dz = w[3 * v.a1 + 2] - w[3 * v.a2 + 2];
        const double dd = sqrt(dx * dx + dy * dy + dz * dz);
        const double val = exp(-c3_k[v.cls] * (dd - c3_r0[v.cls]));
        xv[(long)n * ncap + t] = val;
        if (xvf) xvf[(long)n * ncap + t] = (float)val;
    }
}

// mbpol-3b-poly.cpp:208-312: switch s = sab*sac + sab*sbc + sac*sbc, chain rule, switch
// gradients.  Record: E (kcal/mol) and dE/dx of the 9 atoms; block partials as k_finish2.
__global__ __launch_bounds__(128) void k_finish3(const int* __restrict__ ncnt, int ncap, int nsl, const double* __restrict__ clus,
                                                 const double* __restrict__ xv, const double* __restrict__ part,
                                                 double* __restrict__ erec, double* __restrict__ grec,
                                                 double* __restrict__ blk, const int* __restrict__ trimers,
                                                 const int* __restrict__ wbase, const double* __restrict__ x,
                                                 const double* __restrict__ y, const double* __restrict__ z,
                                                 double* __restrict__ srec)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = min(*ncnt, ncap);
    double red[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (t < nt && trimers) {
        double S[3];
        cluster_raw_sum<3>(trimers + 3L * t, wbase, x, y, z, S);
        srec[t] = S[0];
        srec[(long)ncap + t] = S[1];
        srec[2L * ncap + t] = S[2];
    }
    if (t < nt) {
        double w[27];
#pragma unroll
        for (int k = 0; k < 27; ++k) w[k] = clus[(long)k * ncap + t];
        double E = 0.0;
        for (int q = 0; q < nsl; ++q) E += part[((long)q * 37) * ncap + t];
        double rab[3], rac[3], rbc[3];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            rab[n] = w[n] - w[9 + n];
            rac[n] = w[n] - w[18 + n];
            rbc[n] = w[9 + n] - w[18 + n];
        }
        const double drab = sqrt(rab[0] * rab[0] + rab[1] * rab[1] + rab[2] * rab[2]);
        const double drac = sqrt(rac[0] * rac[0] + rac[1] * rac[1] + rac[2] * rac[2]);
        const double drbc = sqrt(rbc[0] * rbc[0] + rbc[1] * rbc[1] + rbc[2] * rbc[2]);
        double gab, gac, gbc;
        const double sab = f_switch(drab, 0.0, MBGPU_R3F, gab);
        const double sac = f_switch(drac, 0.0, MBGPU_R3F, gac);
        const double sbc = f_switch(drbc, 0.0, MBGPU_R3F, gbc);
        const double sw = sab * sac + sab * sbc + sac * sbc;
        double gc[27];
#pragma unroll
        for (int k = 0; k < 27; ++k) gc[k] = 0.0;
#pragma unroll
        for (int n = 0; n < 36; ++n) {
            double gn = 0.0;
            for (int q = 0; q < nsl; ++q) gn += part[((long)q * 37 + 1 + n) * ncap + t];
            gn *= sw;
            const V3Def v = c3_vars[n];
            const double dx[3] = {w[3 * v.a1] - w[3 * v.a2], w[3 * v.a1 + 1] - w[3 * v.a2 + 1],
                                  w[3 * v.a1 + 2] - w[3 * v.a2 + 2]};
            const double dd = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            const double k = c3_k[v.cls];
            const double gg = -k * gn * xv[(long)n * ncap + t] / dd;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                gc[3 * v.a1 + i] += gg * dx[i];
                gc[3 * v.a2 + i] -= gg * dx[i];
            }
        }
        gab *= (sac + sbc) * E / drab;
        gac *= (sab + sbc) * E / drac;
        gbc *= (sab + sac) * E / drbc;
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            gc[n] += gab * rab[n] + gac * rac[n];
            gc[9 + n] += gbc * rbc[n] - gab * rab[n];
            gc[18 + n] -= gac * rac[n] + gbc * rbc[n];
        }
        const double Eout = E * sw;
        erec[t] = Eout;
#pragma unroll
        for (int k = 0; k < 27; ++k) grec[(long)k * ncap + t] = gc[k];
        red[0] = Eout * MBGPU_ENGUNIT;
        cluster_virial<27>(w, gc, red);
    }
    block_reduce_store(red, 8, blk + 8 * blockIdx.x);
}

// position of value v in the ascending list L[0..n) (v is known to be present)
__device__ __forceinline__ int find_rank(const int* __restrict__ L, int n, int v)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (L[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Per-water gather of the cluster records (replaces update_energy_mbpol / fxx updates /
// update_forces of do_dimer and do_trimer).  One block of 256 threads per water w; the
// clusters of w are found from the lists, without atomics on global memory:
//   A  dimers (w, j), j > w:  off2[w] + rank of j in nbr2[w]         (contiguous)
//   B  dimers (j, w), j < w:  off2[j] + rank of w in nbr2[j]          (j from list 2l)
//   C  trimers centred on w:  toff[w] .. toff[w] + tcnt[w]            (contiguous)
//   D  trimers with w as a leaf of centre c (c a short neighbour of w): the kept pairs of
//      c that contain w's rank r in c's short list, through kmask / rowoff; listed in LDS
//      first (one thread per centre), then processed like the others.
// Each thread sums its items in 2^-32 fixed point (order-independent, hence
// deterministic) and the block combines the thread sums exactly.  Per atom a of w:
//   f_a   += -g_a * engunit
//   e_a   += E * engunit / k                          (k = 6 dimer, 9 trimer)
//   W_a   += F_a (x) (k x_a - S_c),  F_a = -(g_a / k) engunit,  S_c = sum of the raw
//            positions of the cluster's atoms, i.e. F_a (x) sum_{b != a} (x_a - x_b) as
//            compute_heat_flux2 contracts force_matrix with un-imaged positions.
struct GatherArgs {
    int nw, nd, nt, cap2;
    const int *wbase, *dimers, *trimers;
    const int *nbr2, *cnt2, *off2, *nbr2l, *cnt2l;
    const int *nbr3, *cnt3, *toff, *tcnt, *rowoff;
    const unsigned long long* kmask;
    const double *e2, *g2, *s2, *e3, *g3, *s3, *x, *y, *z;
    double *fx, *fy, *fz, *fout, *ae, *hw;
};

#define GATHER_THREADS 256
#define GATHER_MAXLEAF 4096   // >= 64 centres x 63 pairs: cannot overflow with 64-entry short lists

__global__ __launch_bounds__(GATHER_THREADS) void k_gather_water(GatherArgs A)
{
    __shared__ int leaf[GATHER_MAXLEAF];
    __shared__ int nleaf;
    __shared__ long long red[GATHER_THREADS / 64][74];
    const int w = blockIdx.x;
    const int tid = threadIdx.x;
    if (tid == 0) nleaf = 0;
    __syncthreads();
    // D: list the leaf trimers, one thread per centre
    {
        const int n3 = min(A.cnt3[w], 64);
        if (tid < n3) {
            const int c = A.nbr3[(long)w * 64 + tid];
            const int nc3 = min(A.cnt3[c], 64);
            const int r = find_rank(A.nbr3 + (long)c * 64, nc3, w);
            const unsigned long long* km = A.kmask + (long)c * 64;
            const int* ro = A.rowoff + (long)c * 64;
            const int base = A.toff[c];
            const int k0 = __popcll(km[r]);
            const unsigned long long bit = 1ull << r, lower = bit - 1ull;
            int cnt = k0;
            for (int pp = 0; pp < r; ++pp) cnt += (km[pp] & bit) ? 1 : 0;
            int pos = atomicAdd(&nleaf, cnt);
            for (int k = 0; k < k0; ++k)
                if (pos < GATHER_MAXLEAF) leaf[pos++] = base + ro[r] + k;
            for (int pp = 0; pp < r; ++pp) {
                const unsigned long long mp = km[pp];
                if ((mp & bit) && pos < GATHER_MAXLEAF) leaf[pos++] = base + ro[pp] + __popcll(mp & lower);
            }
        }
    }
    __syncthreads();
    const int o = A.wbase[w];
    double xa[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        xa[a][0] = A.x[o + a];
        xa[a][1] = A.y[o + a];
        xa[a][2] = A.z[o + a];
    }
    long long hi[37], lo[37];
#pragma unroll
    for (int k = 0; k < 37; ++k) hi[k] = lo[k] = 0;
    auto acc = [&](int k, double v) {
        const long long q = __double2ll_rn(v * MBGPU_FX_SCALE);
        hi[k] += q >> 32;
        lo[k] += q & 0xffffffffll;
    };
    const int nA = A.cnt2[w], nB = A.cnt2l[w], nC = A.tcnt[w], nD = min(nleaf, GATHER_MAXLEAF);
    const int ntot = nA + nB + nC + nD;
    for (int it = tid; it < ntot; it += GATHER_THREADS) {
        bool tri;
        long c;
        int slot;
        if (it < nA) {
            tri = false;
            c = (long)A.off2[w] + it;
            slot = 0;
        } else if (it < nA + nB) {
            const int j = A.nbr2l[(long)w * A.cap2 + (it - nA)];
            tri = false;
            c = (long)A.off2[j] + find_rank(A.nbr2 + (long)j * A.cap2, A.cnt2[j], w);
            slot = 1;
        } else {
            tri = true;
            c = it < nA + nB + nC ? (long)A.toff[w] + (it - nA - nB) : (long)leaf[it - nA - nB - nC];
            const int* tr = A.trimers + 3 * c;
            slot = tr[0] == w ? 0 : (tr[1] == w ? 1 : 2);
        }
        const double frac = tri ? 1.0 / 9.0 : 1.0 / 6.0;
        const double kk = tri ? 9.0 : 6.0;
        const long nc = tri ? A.nt : A.nd;
        const double E = tri ? A.e3[c] : A.e2[c];
        const double* G = tri ? A.g3 : A.g2;
        const double* Sp = tri ? A.s3 : A.s2;
        const double S[3] = {Sp[c], Sp[nc + c], Sp[2 * nc + c]};
        acc(9, frac * E * MBGPU_ENGUNIT);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double fa[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double ga = G[(long)(9 * slot + 3 * a + q) * nc + c];
                acc(3 * a + q, -(ga * MBGPU_ENGUNIT));
                fa[q] = -(frac * ga) * MBGPU_ENGUNIT;
            }
            const double rr[3] = {kk * xa[a][0] - S[0], kk * xa[a][1] - S[1], kk * xa[a][2] - S[2]};
#pragma unroll
            for (int pp = 0; pp < 3; ++pp)
#pragma unroll
                for (int q = 0; q < 3; ++q) acc(10 + 9 * a + 3 * pp + q, fa[pp] * rr[q]);
        }
    }
    // exact combination: wavefront shuffles, then the wavefronts through LDS
    const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
    for (int k = 0; k < 37; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            hi[k] += __shfl_down(hi[k], off, 64);
            lo[k] += __shfl_down(lo[k], off, 64);
        }
        if (lane == 0) {
            red[wid][k] = hi[k];
            red[wid][37 + k] = lo[k];
        }
    }
    __syncthreads();
    if (tid < 37) {
        long long h = 0, l = 0;
        for (int q = 0; q < GATHER_THREADS / 64; ++q) {
            h += red[q][tid];
            l += red[q][37 + tid];
        }
        const double v = (double)h + (double)l * (1.0 / MBGPU_FX_SCALE);
        if (tid < 9) {
            const int i = o + tid / 3, q = tid % 3;
            A.fout[3 * i + q] = v;
            if (A.fx) (q == 0 ? A.fx : (q == 1 ? A.fy : A.fz))[i] += v;
        } else if (tid == 9) {
            A.ae[o] = A.ae[o + 1] = A.ae[o + 2] = v;
        } else {
            const int k = tid - 10, a = k / 9;
            A.hw[9L * (o + a) + (k % 9)] = v;
        }
    }
}

// Sum the nsl slice partials of every (variable, cluster) into slice 0, in slice order:
// n = rows x ncap elements, clusters at or past the count skipped.
__global__ void k_slice_sum(long n, int nsl, const int* __restrict__ ncnt, int ncap, double* __restrict__ part)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n || i % ncap >= *ncnt) return;
    double s = part[i];
    for (int q = 1; q < nsl; ++q) s += part[q * n + i];
    part[i] = s;
}

// FP64 sum of the nsl FP32 slice partials of every (variable, cluster), in slice order.
__global__ void k_slice_sum_f(long n, int nsl, const int* __restrict__ ncnt, int ncap, const float* __restrict__ pf,
                              double* __restrict__ part)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n || i % ncap >= *ncnt) return;
    double s = (double)pf[i];
    for (int q = 1; q < nsl; ++q) s += (double)pf[q * n + i];
    part[i] = s;
}

// Zero the per-atom outputs of the sites that are not water atoms (M sites).
__global__ void k_zero_other_sites(int natms, const unsigned char* __restrict__ is_water, double* __restrict__ fout,
                                   double* __restrict__ ae, double* __restrict__ hw)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natms || is_water[i]) return;
    fout[3 * i] = fout[3 * i + 1] = fout[3 * i + 2] = 0.0;
    ae[i] = 0.0;
    for (int k = 0; k < 9; ++k) hw[9L * i + k] = 0.0;
}

__global__ void k_mark_water(int nw, const int* __restrict__ wbase, unsigned char* __restrict__ is_water)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const int o = wbase[w];
    is_water[o] = is_water[o + 1] = is_water[o + 2] = 1;
}

// Fixed-order sum of nb block partials (8 values each): one wavefront per value.
__global__ __launch_bounds__(512) void k_reduce_blocks(const double* __restrict__ blk, int nb, double* __restrict__ out)
{
    const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double s = 0.0;
    for (int b = lane; b < nb; b += 64) s += blk[8L * b + q];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) out[q] = s;
}

// Heat flux (compute_heat_flux2, heatcurrent.f90:703-728) from per-atom quantities:
//   J_kin = sum_i (e_kin,i + e_pot,i) v_i
//   J_pot,b = sum_i sum_a v_i,a W_i[a][b] - (sigma_ew1p,i . v_i)_b (upper triangle of sigma)
// over the real atoms i < nreal.  Per-block partials, reduced by k_reduce_blocks.
__global__ __launch_bounds__(256) void k_heat_flux(int nreal, const double* __restrict__ vx,
                                                   const double* __restrict__ vy, const double* __restrict__ vz,
                                                   const double* __restrict__ ekin, const double* __restrict__ epot,
                                                   const double* __restrict__ epot2, const double* __restrict__ w,
                                                   const double* __restrict__ w2, const double* __restrict__ sew,
                                                   double* __restrict__ blk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double red[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (i < nreal) {
        const double v[3] = {vx[i], vy[i], vz[i]};
        double e = ekin[i] + epot[i];
        if (epot2) e += epot2[i];
        for (int a = 0; a < 3; ++a) red[a] = e * v[a];
        for (int b = 0; b < 3; ++b) {
            double s = 0.0;
            for (int a = 0; a < 3; ++a) {
                s += v[a] * w[9 * i + 3 * a + b];
                if (w2) s += v[a] * w2[9 * i + 3 * a + b];
            }
            red[3 + b] = s;
        }
        if (sew) {
            const double* sg = sew + 9 * i;
            red[3] -= sg[0] * v[0] + sg[1] * v[1] + sg[2] * v[2];
            red[4] -= sg[1] * v[0] + sg[4] * v[1] + sg[5] * v[2];
            red[5] -= sg[2] * v[0] + sg[5] * v[1] + sg[8] * v[2];
        }
    }
    block_reduce_store(red, 8, blk + 8 * blockIdx.x);
}

// ============================================================ host-side launchers ====

int mb_launch_lists(MbBuffers& B, const CellGeom& g, hipStream_t s)
{
    const int nw = B.nw;
    hipLaunchKernelGGL(k_zero_i, dim3(1), dim3(64), 0, s, B.counts, 4);
    hipLaunchKernelGGL(k_image_monomers, dim3((nw + 127) / 128), dim3(128), 0, s, nw, B.natms, B.list, B.x, B.y,
                       B.z, g, B.mon, B.wbase, B.counts + 2);
    if (B.cells.V.on) {
        const int rc = cells_build(B.cells, nw, B.mon, B.mon + 1, B.mon + 2, 9, s);   // the oxygens
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_neighbours, dim3(nw), dim3(64), 0, s, nw, B.mon, g, B.cells.V, B.cap2, B.nbr2, B.cnt2,
                       B.nbr2l, B.cnt2l, B.cap3, B.nbr3, B.cnt3, B.counts + 2);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, B.cnt2, nw, B.off2, B.counts + 0);
    hipLaunchKernelGGL(k_fill_dimers, dim3((nw + 127) / 128), dim3(128), 0, s, nw, B.cap2, B.nbr2, B.cnt2, B.off2,
                       B.dimers, B.dcap);
    hipLaunchKernelGGL(k_trimers<false>, dim3(nw), dim3(64), 0, s, nw, B.mon, g, B.nbr3, B.cnt3, B.tcnt, B.toff,
                       B.kmask, B.rowoff, B.trimers, B.tcap);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, B.tcnt, nw, B.toff, B.counts + 1);
    hipLaunchKernelGGL(k_trimers<true>, dim3(nw), dim3(64), 0, s, nw, B.mon, g, B.nbr3, B.cnt3, B.tcnt, B.toff,
                       B.kmask, B.rowoff, B.trimers, B.tcap);
    return (int)hipGetLastError();
}

// Cluster pipeline of the dimers: ncnt (device) dimers of a buffer laid out with stride and
// grid ncap, so that the launches do not depend on the count (a graph may hold them); block
// partials of (energy, virial, stress) go to B.blk[0 .. 8*ceil(ncap/128)), zero for the
// blocks past the count.
int mb_launch_dimers(MbBuffers& B, const int* ncnt, int ncap, const CellGeom& g, bool gather, hipStream_t s)
{
    if (ncap <= 0) return 0;
    const int nb = (ncap + 127) / 128;
    if (gather) hipLaunchKernelGGL(k_gather2, dim3(nb), dim3(128), 0, s, ncnt, ncap, B.dimers, B.mon, g, B.clus2);
    const long n = 32L * ncap;
    if (B.mixed & 1) {
        hipLaunchKernelGGL(k_setup2, dim3(nb), dim3(128), 0, s, ncnt, ncap, B.clus2, B.xv2, B.xv2f);
        launch_pip2b_f32(B.xv2f, B.part2f, ncnt, ncap, B.cs2f, s);
        hipLaunchKernelGGL(k_slice_sum_f, dim3((n + 255) / 256), dim3(256), 0, s, n, pip2b_nslices(), ncnt, ncap,
                           B.part2f, B.part2);
    } else {
        hipLaunchKernelGGL(k_setup2, dim3(nb), dim3(128), 0, s, ncnt, ncap, B.clus2, B.xv2, nullptr);
        launch_pip2b_f64(B.xv2, B.part2, ncnt, ncap, B.cs2, s);
        hipLaunchKernelGGL(k_slice_sum, dim3((n + 255) / 256), dim3(256), 0, s, n, pip2b_nslices(), ncnt, ncap,
                           B.part2);
    }
    hipLaunchKernelGGL(k_finish2, dim3(nb), dim3(128), 0, s, ncnt, ncap, 1, B.clus2, B.xv2, B.part2, B.e2, B.g2,
                       B.blk, gather ? B.dimers : nullptr, B.wbase, B.x, B.y, B.z, B.s2);
    return (int)hipGetLastError();
}

// Same for trimers; block partials after blk_off blocks of the dimers'.
int mb_launch_trimers(MbBuffers& B, const int* ncnt, int ncap, int blk_off, const CellGeom& g, bool gather,
                      hipStream_t s)
{
    if (ncap <= 0) return 0;
    const int nb = (ncap + 127) / 128;
    if (gather) hipLaunchKernelGGL(k_gather3, dim3(nb), dim3(128), 0, s, ncnt, ncap, B.trimers, B.mon, g, B.clus3);
    const long n = 37L * ncap;
    if (B.mixed & 2) {
        hipLaunchKernelGGL(k_setup3, dim3(nb), dim3(128), 0, s, ncnt, ncap, B.clus3, B.xv3, B.xv3f);
        launch_pip3b_f32(B.xv3f, B.part3f, ncnt, ncap, B.cs3f, s);
        hipLaunchKernelGGL(k_slice_sum_f, dim3((n + 255) / 256), dim3(256), 0, s, n, pip3b_nslices(), ncnt, ncap,
                           B.part3f, B.part3);
    } else {
        hipLaunchKernelGGL(k_setup3, dim3(nb), dim3(128), 0, s, ncnt, ncap, B.clus3, B.xv3, nullptr);
        launch_pip3b_f64(B.xv3, B.part3, ncnt, ncap, B.cs3, s);
        hipLaunchKernelGGL(k_slice_sum, dim3((n + 255) / 256), dim3(256), 0, s, n, pip3b_nslices(), ncnt, ncap,
                           B.part3);
    }
    hipLaunchKernelGGL(k_finish3, dim3(nb), dim3(128), 0, s, ncnt, ncap, 1, B.clus3, B.xv3, B.part3, B.e3, B.g3,
                       B.blk + 8L * blk_off, gather ? B.trimers : nullptr, B.wbase, B.x, B.y, B.z, B.s3);
    return (int)hipGetLastError();
}

// Per-water gather of the cluster records (the lists' dimers and trimers, strides dcap and
// tcap) into per-atom forces, energies and heat-flux tensors; scal[0..7] = 2B (energy,
// virial, stress 6), scal[8..15] = 3B.
int mb_launch_accumulate(MbBuffers& B, double* fx, double* fy, double* fz, hipStream_t s)
{
    const int na = B.natms, nw = B.nw;
    const int nb2 = (int)((B.dcap + 127) / 128), nb3 = (int)((B.tcap + 127) / 128);
    if (nb2 + nb3 > B.nblk_cap) return -1;
    hipLaunchKernelGGL(k_zero_i, dim3((na + 1023) / 1024), dim3(256), 0, s, reinterpret_cast<int*>(B.is_water),
                       (na + 3) / 4);
    hipLaunchKernelGGL(k_mark_water, dim3((nw + 255) / 256), dim3(256), 0, s, nw, B.wbase, B.is_water);
    hipLaunchKernelGGL(k_zero_other_sites, dim3((na + 255) / 256), dim3(256), 0, s, na, B.is_water, B.fout,
                       B.atomic_energy, B.heat_w);
    GatherArgs A;
    A.nw = nw; A.nd = (int)B.dcap; A.nt = (int)B.tcap; A.cap2 = B.cap2;
    A.wbase = B.wbase; A.dimers = B.dimers; A.trimers = B.trimers;
    A.nbr2 = B.nbr2; A.cnt2 = B.cnt2; A.off2 = B.off2; A.nbr2l = B.nbr2l; A.cnt2l = B.cnt2l;
    A.nbr3 = B.nbr3; A.cnt3 = B.cnt3; A.toff = B.toff; A.tcnt = B.tcnt; A.rowoff = B.rowoff; A.kmask = B.kmask;
    A.e2 = B.e2; A.g2 = B.g2; A.s2 = B.s2; A.e3 = B.e3; A.g3 = B.g3; A.s3 = B.s3;
    A.x = B.x; A.y = B.y; A.z = B.z;
    A.fx = fx; A.fy = fy; A.fz = fz; A.fout = B.fout; A.ae = B.atomic_energy; A.hw = B.heat_w;
    hipLaunchKernelGGL(k_gather_water, dim3(nw), dim3(GATHER_THREADS), 0, s, A);
    hipLaunchKernelGGL(k_zero_i, dim3(1), dim3(64), 0, s, reinterpret_cast<int*>(B.scal), 32);
    if (nb2) hipLaunchKernelGGL(k_reduce_blocks, dim3(1), dim3(512), 0, s, B.blk, nb2, B.scal);
    if (nb3) hipLaunchKernelGGL(k_reduce_blocks, dim3(1), dim3(512), 0, s, B.blk + 8L * nb2, nb3, B.scal + 8);
    return (int)hipGetLastError();
}

int mb_launch_heat_flux(int nreal, const double* vx, const double* vy, const double* vz, const double* ekin,
                        const double* epot, const double* epot2, const double* w, const double* w2,
                        const double* sew, double* blk, int nblk_cap, double* out, hipStream_t s)
{
    const int nb = (nreal + 255) / 256;
    if (nb > nblk_cap) return -1;
    hipLaunchKernelGGL(k_heat_flux, dim3(nb), dim3(256), 0, s, nreal, vx, vy, vz, ekin, epot, epot2, w, w2, sew,
                       blk);
    hipLaunchKernelGGL(k_reduce_blocks, dim3(1), dim3(512), 0, s, blk, nb, out);
    return (int)hipGetLastError();
}
