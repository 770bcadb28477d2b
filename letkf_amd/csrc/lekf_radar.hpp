// lekf_radar.hpp -- SCALE radar reflectivity / Doppler velocity obs operator.
#pragma once
#include <hip/hip_runtime.h>

// scale/common/common_obs_scale.f90:71-72,141-151
#define LEKF_ID_RADAR_REF 4001
#define LEKF_ID_RADAR_VR 4002
#define LEKF_IQC_GOOD 0
#define LEKF_IQC_GROSS_ERR 5
#define LEKF_IQC_REF_MEM 12
#define LEKF_IQC_RADAR_VHI 19
#define LEKF_IQC_OUT_VHI 20
#define LEKF_IQC_OUT_VLO 21
#define LEKF_IQC_OUT_H 22
#define LEKF_IQC_OBS_BAD 50
#define LEKF_IQC_OTYPE 90
#define LEKF_ID_RADAR_PRH 4003
#define LEKF_ID_H08IR 8800
#define LEKF_ID_SCALE_RAIN 19999
#define LEKF_ID_TCLON 99991
#define LEKF_ID_TCLAT 99992
#define LEKF_ID_TCMIP 99993
#define LEKF_UNDEF (-9.99e33)

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

// SCALE's departure and QC step of set_letkf_obs, one thread per obs.
struct ScaleQcArgs {
    int nobs, ne;
    int use_ref, use_vr, use_obserr_ref, use_obserr_vr;
    int min_member, min_member_obsref;
    double thres;          // RADAR_REF_THRES_DBZ, raised to MIN_RADAR_REF_DBZ
    double min_ref_dbz, low_ref_shift;
    double min_ref;        // 10^(MIN_RADAR_REF_DBZ / 10)
    double obserr_ref, obserr_vr;
    double gross, gross_rain, gross_ref, gross_vr, gross_prh, gross_tcx, gross_tcy, gross_tcp;
    const double *elm, *dat_in, *err_in;   // nobs each
    long so, sm;           // member m of obs n at [n * so + m * sm] in qc_mem and hx
    const int* qc_mem;     // the operator's qc of each member, or null
    double* hx;            // H(x) in, perturbations out where qc is good
    double *dat, *err, *dep;
    int* qc;
};

hipError_t lekf_scale_qc_launch(const ScaleQcArgs& A, hipStream_t st);
