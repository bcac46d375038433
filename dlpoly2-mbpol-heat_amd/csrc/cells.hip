// cells.hip -- link cells for the neighbour searches (electrostatics rows, dispersion rows).
//
// The reference finds pairs through Verlet lists built over link cells (parlink.f): the box
// is cut into nc_a cells per axis, each at least the cut-off wide, and a particle's
// partners lie in its own and the 26 neighbouring cells.  Here the cells are laid out in
// fractional coordinates s = (a*_a . r) wrapped into [0, 1), so every imcon (cubic,
// orthorhombic, triclinic) is handled alike: a pair within the cut-off under the minimum
// image has |ds_a| < cutoff / width_a <= 1 / nc_a in each axis, so the two cells differ by at
// most one in each index (mod nc_a).  With fewer than three cells on an axis the searches
// scan every particle instead.
//
// Build (one stream, no host round trip): k_cell_keys gives each particle its cell,
// k_cell_sort (one workgroup) counts, scans and places the particles by cell, and
// k_cell_order puts each cell's particles in ascending index.  Search order is then fixed
// by the positions alone: the 27 cells in a fixed order, ascending index within a cell.
#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <string>

#include "mbgpu_internal.h"

namespace {

__global__ __launch_bounds__(256) void k_cell_keys(CellView V, int n, const double* __restrict__ x,
                                                   const double* __restrict__ y, const double* __restrict__ z,
                                                   int stride, int* __restrict__ key)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const long o = (long)i * stride;
    int c[3];
    for (int a = 0; a < 3; ++a) {
        double s = V.rv[a][0] * x[o] + V.rv[a][1] * y[o] + V.rv[a][2] * z[o];
        s -= floor(s);
        int k = (int)(s * V.nc[a]);
        // non-finite positions land in cell 0: the kernels' own distance tests decide
        c[a] = (k >= 0 && k < V.nc[a]) ? k : 0;
    }
    key[i] = (c[0] * V.nc[1] + c[1]) * V.nc[2] + c[2];
}

// Counting sort of the particles by cell in one workgroup: per-cell counts in LDS, their
// exclusive scan, then each particle placed at its cell's next free slot (LDS atomics: the
// order within a cell is arbitrary here and fixed by k_cell_order).  start[ncell] = n.
__global__ __launch_bounds__(1024) void k_cell_sort(int ncell, int n, const int* __restrict__ key,
                                                    int* __restrict__ start, int* __restrict__ items)
{
    __shared__ int cnt[CELLS_MAX_CELLS];
    __shared__ int wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int c = tid; c < ncell; c += 1024) cnt[c] = 0;
    __syncthreads();
    for (int i = tid; i < n; i += 1024) atomicAdd(&cnt[key[i]], 1);
    __syncthreads();
    // exclusive scan of cnt: each thread a contiguous run of cells, then the run totals
    const int per = (ncell + 1023) / 1024, c0 = tid * per, c1 = min(ncell, c0 + per);
    int run = 0;
    for (int c = c0; c < c1; ++c) run += cnt[c];
    int incl = run;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wv; ++w) base += wsum[w];
    int off = base + incl - run;
    for (int c = c0; c < c1; ++c) {
        const int k = cnt[c];
        start[c] = off;
        cnt[c] = off;
        off += k;
    }
    if (tid == 0) start[ncell] = n;
    __syncthreads();
    for (int i = tid; i < n; i += 1024) items[atomicAdd(&cnt[key[i]], 1)] = i;
}

// One wavefront per cell puts its particles in ascending index: the cell's entries are
// staged in LDS and each gets its rank among them (cells of more than 256 particles, rare
// at the cell sizes used, are ordered by lane 0 in place).
__global__ __launch_bounds__(256) void k_cell_order(int ncell, const int* __restrict__ start, int* __restrict__ items)
{
    constexpr int U = 4;
    __shared__ int st[4][64 * U];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = blockIdx.x * 4 + wv;
    if (c >= ncell) return;
    int* row = items + start[c];
    const int len = start[c + 1] - start[c];
    if (len <= 64 * U) {
        int* sr = st[wv];
        int v[U], r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            v[u] = lane + 64 * u < len ? row[lane + 64 * u] : 0;
            if (lane + 64 * u < len) sr[lane + 64 * u] = v[u];
            r[u] = 0;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (int k = 0; k < len; ++k) {
            const int w = sr[k];
#pragma unroll
            for (int u = 0; u < U; ++u) r[u] += w < v[u];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (lane + 64 * u < len) row[r[u]] = v[u];
    } else if (lane == 0) {
        for (int a = 1; a < len; ++a) {
            const int v = row[a];
            int b = a - 1;
            while (b >= 0 && row[b] > v) {
                row[b + 1] = row[b];
                --b;
            }
            row[b + 1] = v;
        }
    }
}

template <typename T>
void cfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

void cells_free(CellList& C)
{
    cfree(C.key);
    cfree(C.sidx);
    cfree(C.start);
    C.n_cap = 0;
    C.ncell_cap = 0;
    C.V.on = 0;
}

int cells_plan(CellList& C, const double cell[9], double cutoff, int n, int min_particles, std::string& err)
{
    C.V.on = 0;
    // MBGPU_CELLS_MIN=m: use the cells from m particles on, whatever the caller's threshold
    // (a test aid: it makes the cell searches run on boxes where the scans are faster)
    if (const char* ev = getenv("MBGPU_CELLS_MIN")) min_particles = std::max(1, atoi(ev));
    const double *a = cell, *b = cell + 3, *cc = cell + 6;
    const double bxc[3] = {b[1] * cc[2] - b[2] * cc[1], b[2] * cc[0] - b[0] * cc[2], b[0] * cc[1] - b[1] * cc[0]};
    const double cxa[3] = {cc[1] * a[2] - cc[2] * a[1], cc[2] * a[0] - cc[0] * a[2], cc[0] * a[1] - cc[1] * a[0]};
    const double axb[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double v = a[0] * bxc[0] + a[1] * bxc[1] + a[2] * bxc[2];
    if (!(fabs(v) > 0.0) || !std::isfinite(v) || !(cutoff > 0.0) || n < min_particles) return 0;
    const double* cr[3] = {bxc, cxa, axb};
    int nc[3];
    for (int k = 0; k < 3; ++k) {
        const double len = sqrt(cr[k][0] * cr[k][0] + cr[k][1] * cr[k][1] + cr[k][2] * cr[k][2]);
        const double width = fabs(v) / len;
        nc[k] = std::min(CELLS_MAX_PER_AXIS, (int)floor(width / (cutoff * (1.0 + 1e-9))));   // <= 16^3 cells
        if (nc[k] < 3) return 0;   // scan every particle
        for (int d = 0; d < 3; ++d) C.V.rv[k][d] = cr[k][d] / v;
        C.V.nc[k] = nc[k];
    }
    const int ncell = nc[0] * nc[1] * nc[2];
    if (C.n_cap < n || C.ncell_cap < ncell) {
        cells_free(C);
        const size_t nn = (size_t)std::max(n, 1);
        if (mbgpu_alloc(&C.key, nn) || mbgpu_alloc(&C.sidx, nn) || mbgpu_alloc(&C.start, (size_t)ncell + 1)) {
            err = "link cells: device allocation failed";
            return -1;
        }
        C.n_cap = n;
        C.ncell_cap = ncell;
    }
    C.ncell = ncell;
    C.V.key = C.key;
    C.V.idx = C.sidx;
    C.V.start = C.start;
    C.V.on = 1;
    return 0;
}

int cells_build(CellList& C, int n, const double* x, const double* y, const double* z, int stride, hipStream_t s)
{
    if (!C.V.on || n <= 0) return 0;
    hipLaunchKernelGGL(k_cell_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, C.V, n, x, y, z, stride,
                       C.key);
    hipLaunchKernelGGL(k_cell_sort, dim3(1), dim3(1024), 0, s, C.ncell, n, C.key, C.start, C.sidx);
    hipLaunchKernelGGL(k_cell_order, dim3((unsigned)((C.ncell + 3) / 4)), dim3(256), 0, s, C.ncell, C.start, C.sidx);
    return (int)hipGetLastError();
}
