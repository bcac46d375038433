// mbgpu_dev.h -- device-side helpers shared by the MB-pol kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MBGPU_ENGUNIT 418.4             // kcal/mol -> 10 J/mol (mbpol.f90:37)
#define MBGPU_R2F 6.5                   // 2B cutoff (mbpol-2b-poly.cpp:1863)
#define MBGPU_R3F 4.5                   // 3B cutoff (mbpol-3b-poly.cpp:1489)

// Per-atom sums are accumulated in fixed point so that the result does not depend on the
// order in which clusters finish (integer addition is associative).  A contribution v is
// rounded to q = nint(v * 2^32) (|v| < 2^31) and q is split into its high and low 32-bit
// halves, which go to two int64 accumulators: hi collects q >> 32, lo collects q & (2^32-1).
// The sum is hi + lo * 2^-32 exactly, with 2^-32 resolution and no overflow of either
// accumulator for any realistic number of contributions.
#define MBGPU_FX_SCALE 4294967296.0     // 2^32

// 0-based atom offset of water w's oxygen from listttm2 (1-based atom ids, O, H1, H2 per
// water, read at mbpol.f90:184-188), or -1 when w's three entries are not the consecutive
// ids a, a+1, a+2 with 1 <= a and a+2 <= natms.  The kernels address a water's atoms as
// offset + 0..2, as the reference does (o1, o1+1, o1+2, mbpol.f90:229,526-544), so every
// entry is checked by this one predicate, on the host for host arrays and on the device
// for device arrays, before any atom array is indexed with it.
__host__ __device__ inline int water_base(const int* list, int w, int natms)
{
    const int a = list[3 * w], b = list[3 * w + 1], c = list[3 * w + 2];
    return (a >= 1 && a <= natms - 2 && b == a + 1 && c == a + 2) ? a - 1 : -1;
}

struct CellGeom {
    int imcon;          // 0 none (an isolated cluster), 1 cubic, 2 orthorhombic, 3 general
                        // (mbpol.f90:1319-1357)
    double c[9];        // cell vectors, Fortran cell(1..9)
    double rc[9];       // inverse (imcon 3)
    double r[3];        // 1/cell(1), 1/cell(5), 1/cell(9)
};

// Link cells of one set of particles (cells.hip): nc[a] cells per axis in fractional
// coordinates s_a = rv[a] . r wrapped into [0, 1); key[i] the cell of particle i, idx the
// particles by cell (ascending index within a cell), start[c] where cell c begins.
// on = 0: no cells (a search scans every particle).
struct CellView {
    int on;
    int nc[3];
    double rv[3][3];
    const int* key;
    const int* idx;
    const int* start;
};

// f(first, end) for the ranges of idx of the 27 cells around particle i's cell, in a fixed
// order (needs nc >= 3 on every axis, so that the 27 cells are distinct)
template <class F>
__device__ __forceinline__ void for_neighbour_cells(const CellView& V, int i, F&& f)
{
    const int k = V.key[i];
    const int c2 = k % V.nc[2], c1 = (k / V.nc[2]) % V.nc[1], c0 = k / (V.nc[1] * V.nc[2]);
    for (int d0 = -1; d0 <= 1; ++d0) {
        int a0 = c0 + d0;
        a0 += a0 < 0 ? V.nc[0] : 0;
        a0 -= a0 >= V.nc[0] ? V.nc[0] : 0;
        for (int d1 = -1; d1 <= 1; ++d1) {
            int a1 = c1 + d1;
            a1 += a1 < 0 ? V.nc[1] : 0;
            a1 -= a1 >= V.nc[1] ? V.nc[1] : 0;
            for (int d2 = -1; d2 <= 1; ++d2) {
                int a2 = c2 + d2;
                a2 += a2 < 0 ? V.nc[2] : 0;
                a2 -= a2 >= V.nc[2] ? V.nc[2] : 0;
                const int cell = (a0 * V.nc[1] + a1) * V.nc[2] + a2;
                f(V.start[cell], V.start[cell + 1]);
            }
        }
    }
}

// Minimum image of one vector exactly as image_vectors (mbpol.f90:1305-1361):
// d - L*nint(d*(1/L)), nint rounding half away from zero, no FMA contraction.
__device__ __forceinline__ void image_vec(const CellGeom& g, double& dx, double& dy, double& dz)
{
#pragma clang fp contract(off)
    if (g.imcon == 0) return;   // an isolated cluster: no images
    if (g.imcon == 1) {
        dx = dx - g.c[0] * round(dx * g.r[0]);
        dy = dy - g.c[0] * round(dy * g.r[0]);
        dz = dz - g.c[0] * round(dz * g.r[0]);
    } else if (g.imcon == 2) {
        dx = dx - g.c[0] * round(dx * g.r[0]);
        dy = dy - g.c[4] * round(dy * g.r[1]);
        dz = dz - g.c[8] * round(dz * g.r[2]);
    } else {
        const double ssx = g.rc[0] * dx + g.rc[3] * dy + g.rc[6] * dz;
        const double ssy = g.rc[1] * dx + g.rc[4] * dy + g.rc[7] * dz;
        const double ssz = g.rc[2] * dx + g.rc[5] * dy + g.rc[8] * dz;
        const double xss = ssx - round(ssx), yss = ssy - round(ssy), zss = ssz - round(ssz);
        dx = g.c[0] * xss + g.c[3] * yss + g.c[6] * zss;
        dy = g.c[1] * xss + g.c[4] * yss + g.c[7] * zss;
        dz = g.c[2] * xss + g.c[5] * yss + g.c[8] * zss;
    }
}

// Squared length summed in the reference's order ((x^2 + y^2) + z^2), no contraction.
__device__ __forceinline__ double norm2_nc(double x, double y, double z)
{
#pragma clang fp contract(off)
    const double a = x * x;
    const double b = y * y;
    const double c = z * z;
    return (a + b) + c;
}

// add v to the fixed-point accumulator whose high half is hi[k] and low half lo[k]
__device__ __forceinline__ void fx_add(long long* hi, long long* lo, long k, double v)
{
    const long long q = __double2ll_rn(v * MBGPU_FX_SCALE);
    atomicAdd(reinterpret_cast<unsigned long long*>(hi + k), (unsigned long long)(q >> 32));
    atomicAdd(reinterpret_cast<unsigned long long*>(lo + k), (unsigned long long)(q & 0xffffffffll));
}

__device__ __forceinline__ double fx_value(const long long* hi, const long long* lo, long k)
{
    return (double)hi[k] + (double)lo[k] * (1.0 / MBGPU_FX_SCALE);
}
