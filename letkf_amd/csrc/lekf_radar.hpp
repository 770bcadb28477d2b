// lekf_radar.hpp -- SCALE radar reflectivity / Doppler velocity obs operator.
#pragma once
#include <hip/hip_runtime.h>

// scale/common/common_obs_scale.f90:71-72,141-151
#define LEKF_ID_RADAR_REF 4001
#define LEKF_ID_RADAR_VR 4002
#define LEKF_IQC_GOOD 0
#define LEKF_IQC_RADAR_VHI 19
#define LEKF_IQC_OUT_VHI 20
#define LEKF_IQC_OUT_VLO 21
#define LEKF_IQC_OUT_H 22
#define LEKF_IQC_OTYPE 90

struct RadarArgs {
    int nobs, nx, ny, nlev, ne, nv3d;
    int var[11];           // 3-D variable of u, v, w, t, p, q, qc, qr, qi, qs, qg
    int method, use_wt;    // METHOD_REF_CALC, USE_TERMINAL_VELOCITY
    double min_ref_dbz, low_ref_shift, zmax;
    double gam[3];         // Gamma(4.8), Gamma(4.25), Gamma(4.5)
    const double *elm, *lon, *lat, *z, *ri, *rj, *rlon, *rlat, *rz;  // nobs each
    const double* x3;      // ensemble ((lev * nij + ij) * nv3d + v) * ne + m
    const double* hgt;     // level heights hgt[ij + nij * lev]
    double* hxf;           // [n * ne + m]
    int* qc;               // [n * ne + m]
    double* rk;            // [n]
};

hipError_t lekf_radar_launch(const RadarArgs& A, hipStream_t st);
