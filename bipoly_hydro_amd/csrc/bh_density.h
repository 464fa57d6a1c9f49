// bh_density.h -- device density generator of the served workload
// (bh_density.hip) and the host bipolytrope profile it interpolates.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "bh_ops.h"

namespace bh {

struct DensityArgs {
    double *rho;          // (numphi, numz, numr) output
    const double *prof;   // f(u), u = d / R on [0, 1], nprof samples
    int nprof;
    int numr, numz, numphi;
    double dr, dphi;
    double rho_c[2], R[2], cx[2];  // per star: central density, radius, centre x (physical units)
    double noise;
    unsigned long long seed;
    int isym;  // 2, 3: the grid starts at the equator, z_K = (K - 1.5) dz
};
Op density_op(const DensityArgs &a);

// Normalised density profile rho(u R) / rho_c of a bipolytrope: core of
// index n_core out to core_frac of the radius, envelope of index n_env,
// with density and pressure continuous at the interface.  Tabulated at
// u = i / (ntab - 1).  info[0] = surface radius and info[1] = interface
// radius in the core's Lane-Emden units.
int bipolytrope_profile(double n_core, double n_env, double core_frac, int ntab, std::vector<double> *table,
                        double info[2]);

}  // namespace bh
