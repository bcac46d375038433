// ring.hip -- in-process ring polymer: normal-mode matrices, bead transforms and spring
// forces on the device.
//
// The reference runs one MPI rank group per bead and exchanges coordinates around the ring
// (centroid_matrix.f, centroid_spring_force.f).  Here every bead of a ring polymer lives in
// one device array, bead-major with a bead's coordinates as [3][natom] (the x, y, z blocks
// of a DL_POLY configuration), so the transforms are one small matrix product over the bead
// index and the springs need no exchange:
//   transform:  out[k][c] = sum_b M[k][b] in[b][c]      (centroid_matrix.f:140-476)
//   springs:    for bead k and atom i, with d1 = x_k - xc, d2 = x_k - x_{k-1},
//               d3 = x_k - x_{k+1} (ring order), f_k -= coeff m (d2 + d3) fac3 and the sums
//               sum m |d2|^2 and sum -d1 . f_k (f before the springs)
//               (centroid_spring_force.f:104-142).
// Sums are taken in a fixed order (thread, wavefront, block, then blocks in order).
#include "mbgpu_internal.h"

namespace {

__device__ __forceinline__ double rg_wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_ring_transform(int P, long ncoord, const double* __restrict__ M,
                                                        const double* __restrict__ in, double* __restrict__ out)
{
    const long c = (long)blockIdx.x * 256 + threadIdx.x;
    const int k = blockIdx.y;
    if (c >= ncoord) return;
    double s = 0.0;
    for (int b = 0; b < P; ++b) s += M[k * P + b] * in[(long)b * ncoord + c];
    out[(long)k * ncoord + c] = s;
}

__global__ __launch_bounds__(256) void k_ring_spring(int P, int natom, double coeff, double fac3,
                                                     const double* __restrict__ mass, const double* __restrict__ x,
                                                     const double* __restrict__ xc, double* __restrict__ f,
                                                     double* __restrict__ blk)
{
    __shared__ double red[2][4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int k = blockIdx.y;
    double se = 0.0, sd = 0.0;
    if (i < natom) {
        const long nb = 3L * natom;
        const int kp = (k + P - 1) % P, kn = (k + 1) % P;
        double d1[3], d2[3], d3[3];
        for (int a = 0; a < 3; ++a) {
            const double xk = x[k * nb + (long)a * natom + i];
            d1[a] = xk - xc[(long)a * natom + i];
            d2[a] = xk - x[kp * nb + (long)a * natom + i];
            d3[a] = xk - x[kn * nb + (long)a * natom + i];
        }
        double* fk = f + k * nb;
        const double m = mass[i];
        sd = -d1[0] * fk[i] - d1[1] * fk[natom + i] - d1[2] * fk[2L * natom + i];
        se = m * (d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2]);
        for (int a = 0; a < 3; ++a) fk[(long)a * natom + i] -= coeff * m * (d2[a] + d3[a]) * fac3;
    }
    se = rg_wsum(se);
    sd = rg_wsum(sd);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = se;
        red[1][threadIdx.x >> 6] = sd;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int q = threadIdx.x;
        blk[2L * ((long)k * gridDim.x + blockIdx.x) + q] = ((red[q][0] + red[q][1]) + red[q][2]) + red[q][3];
    }
}

__global__ __launch_bounds__(64) void k_ring_sum(int nb, const double* __restrict__ blk, double* __restrict__ out)
{
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nb; k += 64) {
        a += blk[2 * k];
        b += blk[2 * k + 1];
    }
    a = rg_wsum(a);
    b = rg_wsum(b);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
    }
}

}  // namespace

int ring_launch_transform(int P, long ncoord, const double* M, const double* in, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_ring_transform, dim3((unsigned)((ncoord + 255) / 256), P), dim3(256), 0, s, P, ncoord, M, in,
                       out);
    return (int)hipGetLastError();
}

// blk: 2 * P * ceil(natom / 256) doubles; out: 2 doubles (sum m |d2|^2, sum -d1 . f)
int ring_launch_spring(int P, int natom, double coeff, double fac3, const double* mass, const double* x,
                       const double* xc, double* f, double* blk, double* out, hipStream_t s)
{
    const int nbx = (natom + 255) / 256;
    hipLaunchKernelGGL(k_ring_spring, dim3(nbx, P), dim3(256), 0, s, P, natom, coeff, fac3, mass, x, xc, f, blk);
    hipLaunchKernelGGL(k_ring_sum, dim3(1), dim3(64), 0, s, P * nbx, blk, out);
    return (int)hipGetLastError();
}
