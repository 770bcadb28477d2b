// lekf_relax.hpp -- argument block of the analysis relaxation pass.
#pragma once
#include <hip/hip_runtime.h>

struct RelaxArgs {
    long npt;                   // points (lev * nij + ij)
    int nij, ne, nv3d, nv2d;
    int mode;                   // LEKF_RELAX_NONE / RTPP / RTPS
    double alpha;               // RELAX_ALPHA or RELAX_ALPHA_SPREAD
    double qmax;                // Q_SPRD_MAX (<= 0: off)
    int iq;                     // 3-D row of q
    int nx, ny;                 // regional grid (boundary buffer)
    double dx, dy, width;       // BOUNDARY_BUFFER_WIDTH (<= 0: off)
    const double *x3b, *x2b;
    double *x3a, *x2a;
};

hipError_t lekf_relax_launch(const RelaxArgs& A, hipStream_t st);
