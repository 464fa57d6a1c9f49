// bh_density.hip -- device generator of the served workload's densities: a
// binary of two bipolytropes (core and envelope of different polytropic
// index) on the solver's cylindrical grid, so that a self-consistent-field
// sequence of solves keeps its density on the device.
//
// Grid (setup_frac.F90:407-463): r_J = (J - 1.5) dr, z_K = (K - 0.5 -
// numz/2) dz (isym = 1) or (K - 1.5) dz (isym = 2, 3), phi_L = (L - 1) 2 pi /
// numphi, dz = dr.  Each star
// adds rho_c f(d / R) inside its radius, f tabulated on [0, 1] and
// interpolated linearly, plus noise * rho_c * (2 h - 1), h a splitmix64
// hash of (seed, star, point index) in [0, 1).
#include "bh_density.h"
#include "bh_ops.h"

#include <hip/hip_runtime.h>

namespace bh {

__device__ __forceinline__ double hash_uniform(unsigned long long x)
{
    // splitmix64 finaliser (Steele, Lea, Flood, OOPSLA 2014)
    unsigned long long z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ void density_tile(const DensityArgs &a, int bx, int by, int)
{
    const int numr = a.numr, numz = a.numz;
    const long plane = (long)numr * numz;
    const long t = (long)bx * 256 + threadIdx.x;
    if (t >= plane) return;
    const int L = by;
    const int J = (int)(t % numr) + 1, K = (int)(t / numr) + 1;
    const double r = ((double)J - 1.5) * a.dr;
    const double z = a.isym == 1 ? ((double)K - 0.5 - (double)(numz / 2)) * a.dr : ((double)K - 1.5) * a.dr;
    const double phi = (double)L * a.dphi;
    const double x = r * cos(phi), y = r * sin(phi);
    const long idx = (long)L * plane + t;
    double rho = 0.0;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const double dx = x - a.cx[s];
        const double d = sqrt(dx * dx + y * y + z * z);
        if (d < a.R[s]) {
            const double u = d / a.R[s] * (double)(a.nprof - 1);
            int i = (int)u;
            if (i > a.nprof - 2) i = a.nprof - 2;
            const double w = u - (double)i;
            const double f = a.prof[i] + w * (a.prof[i + 1] - a.prof[i]);
            rho += a.rho_c[s] * f;
            if (a.noise != 0.0) {
                const double h = hash_uniform(a.seed * 1000003ull + (unsigned long long)s * 0x100000000ull +
                                              (unsigned long long)idx);
                rho += (2.0 * h - 1.0) * a.noise * a.rho_c[s];
            }
        }
    }
    a.rho[idx] = rho;
}

Op density_op(const DensityArgs &a)
{
    const long plane = (long)a.numr * a.numz;
    return Op::make(OP_DENSITY, 0, a, (int)((plane + 255) / 256), a.numphi, 1);
}

__global__ __launch_bounds__(256) void density_kernel(DensityArgs a, int nx, int ny, int total)
{
    BH_SINGLE_LOOP(a, nx, ny, total, (density_tile(a, bx, by, bz)));
}

__global__ __launch_bounds__(256) void density_grouped(const DensityArgs *__restrict__ args, GroupSched gs)
{
    BH_GROUP_LOOP(DensityArgs, args, gs, (density_tile(ga, bx, by, bz)));
}

bool density_launch(const Op &op, hipStream_t s)
{
    if (op.kind != OP_DENSITY) return false;
    const int nb = persistent_blocks(op.tiles(), kernel_occupancy((const void *)density_kernel, 256, 0));
    hipLaunchKernelGGL(density_kernel, dim3(nb), dim3(256), 0, s, op.as<DensityArgs>(), op.nx, op.ny,
                       (int)op.tiles());
    return true;
}

bool density_launch_grouped(int kind, int, const void *d_args, const GroupSched &gs, int nblocks, hipStream_t s)
{
    if (kind != OP_DENSITY) return false;
    hipLaunchKernelGGL(density_grouped, dim3(nblocks), dim3(256), 0, s, (const DensityArgs *)d_args, gs);
    return true;
}

bool density_occupancy(int kind, int, int *blocks)
{
    if (kind != OP_DENSITY) return false;
    *blocks = kernel_occupancy((const void *)density_grouped, 256, 0);
    return true;
}

}  // namespace bh
