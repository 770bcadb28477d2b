// lekf_obsope.hpp -- argument block of the global obs-operator kernel.
#pragma once
#include <hip/hip_runtime.h>

struct ObsOpArgs {
    int nobs, ne, nlon, nlat, nlev, nv3d, nv2d;
    int iu, iv, it, iq, ips;        // variable indices (3-D: u, v, t, q; 2-D: ps)
    const double *lat, *sig;        // nlat latitudes (deg), nlev sigma levels
    const double *x3, *x2;          // ensemble rows [(lev*nij + ij)*nv3d + v][m], [(ij*nv2d + v)][m]
    const double *elm, *olon, *olat, *olev, *odat;
    double* hxf;                    // out: H(x) of member m at hxf[n*ne + m] (0 if rejected)
    int* qc;                        // out: 1 if every member accepts the obs
    double *ri, *rj, *rk;           // optional out: grid coordinates (rk per member)
};

hipError_t lekf_obsope_global_launch(const ObsOpArgs& A, hipStream_t st);
