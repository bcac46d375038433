// md.hip -- velocity-Verlet stages on the device for the real atoms of a context.
//
// The reference integrates with DL_POLY's Verlet leapfrog (nve_0.f:50-110: v(t+dt/2) =
// v(t-dt/2) + dt F/m, x(t+dt) = x(t) + dt v(t+dt/2), full-step velocities as the average of
// the two half steps for the kinetic energy).  The velocity-Verlet form used here produces
// the same positions from the same start and has the full-step velocities directly, which
// is what the heat flux needs:
//   stage 1: v += (dt/2) F(t)/m;  x += dt v;  the force arrays are cleared for the next
//            evaluation, which accumulates into them;
//   stage 2: v += (dt/2) F(t+dt)/m;  e_kin,i = m v^2 / 2 and its sum.
// Units are the reference's internal ones (amu, Angstrom, ps, 10 J/mol), in which F/m is an
// acceleration in Angstrom/ps^2.  One thread per real atom; the kinetic-energy sum is
// reduced in a fixed order (block partials, then one block).
#include "mbgpu_internal.h"

namespace {

__device__ __forceinline__ double md_wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_vv_stage(int stage, int nr, double dt, const double* __restrict__ mass,
                                                  double* __restrict__ x, double* __restrict__ y,
                                                  double* __restrict__ z, double* __restrict__ vx,
                                                  double* __restrict__ vy, double* __restrict__ vz,
                                                  double* __restrict__ fx, double* __restrict__ fy,
                                                  double* __restrict__ fz, double* __restrict__ fsite,
                                                  double* __restrict__ ekin, double* __restrict__ blk)
{
    __shared__ double red[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0;
    if (i < nr) {
        double f[3] = {0.0, 0.0, 0.0};
        if (fx) {
            f[0] += fx[i];
            f[1] += fy[i];
            f[2] += fz[i];
        }
        if (fsite) {
            f[0] += fsite[3 * i];
            f[1] += fsite[3 * i + 1];
            f[2] += fsite[3 * i + 2];
        }
        const double h = 0.5 * dt / mass[i];
        const double v0 = vx[i] + h * f[0], v1 = vy[i] + h * f[1], v2 = vz[i] + h * f[2];
        vx[i] = v0;
        vy[i] = v1;
        vz[i] = v2;
        if (stage == 1) {
            x[i] += dt * v0;
            y[i] += dt * v1;
            z[i] += dt * v2;
            if (fx) fx[i] = fy[i] = fz[i] = 0.0;
            if (fsite) fsite[3 * i] = fsite[3 * i + 1] = fsite[3 * i + 2] = 0.0;
        } else {
            ke = 0.5 * mass[i] * (v0 * v0 + v1 * v1 + v2 * v2);
            if (ekin) ekin[i] = ke;
        }
    }
    const double s = md_wsum(ke);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ __launch_bounds__(64) void k_vv_sum(int nb, const double* __restrict__ blk, double* __restrict__ out)
{
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += 64) s += blk[b];
    s = md_wsum(s);
    if (threadIdx.x == 0) out[0] = s;
}

// Heat-current autocorrelation C(tau) = (1 / (nt - tau)) sum_{t < nt - tau} J(t) . J(t + tau)
// for J stored [nt][3]: one block per lag, the threads stride over t; thread sums, then the
// wavefront sums and the four wave sums in order (the result does not depend on timing).
__global__ __launch_bounds__(256) void k_hcacf(long nt, const double* __restrict__ j, double* __restrict__ acf)
{
    __shared__ double red[4];
    const long tau = blockIdx.x;
    const long n = nt - tau;
    double s = 0.0;
    for (long t = threadIdx.x; t < n; t += 256) {
        const double* a = j + 3 * t;
        const double* b = j + 3 * (t + tau);
        s += a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    }
    s = md_wsum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) acf[tau] = (((red[0] + red[1]) + red[2]) + red[3]) / (double)n;
}

}  // namespace

// lags 0..ntau-1 of a series of nt >= ntau vectors
int md_launch_hcacf(long nt, int ntau, const double* j, double* acf, hipStream_t s)
{
    hipLaunchKernelGGL(k_hcacf, dim3(ntau), dim3(256), 0, s, nt, j, acf);
    return (int)hipGetLastError();
}

// nothing outside the nr real atoms is touched; blk needs ceil(nr/256) doubles
int md_launch_stage(int stage, int nr, double dt, const double* mass, double* x, double* y, double* z, double* vx,
                    double* vy, double* vz, double* fx, double* fy, double* fz, double* fsite, double* ekin,
                    double* blk, double* out, hipStream_t s)
{
    const int nb = (nr + 255) / 256;
    hipLaunchKernelGGL(k_vv_stage, dim3(nb), dim3(256), 0, s, stage, nr, dt, mass, x, y, z, vx, vy, vz, fx, fy, fz,
                       fsite, ekin, blk);
    hipLaunchKernelGGL(k_vv_sum, dim3(1), dim3(64), 0, s, nb, blk, out);
    return (int)hipGetLastError();
}
