// lekf_cull.hpp -- argument blocks of the selection / obs-index kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "lekf_types.h"

#define LEKF_MAX_LEV 128    // levels per column in the global cull
#define LEKF_MAX_BANDS 512  // latitude bands per rectangle
#define LEKF_TYP_PS 5       // obs type index of surface pressure (u,v,t,q,rh,ps,rain)

struct GlobalCullArgs {
    int nij, nlev, nlon, nlat;
    const double* lon;           // nlon grid longitudes (deg)
    const double* lat;           // nlat grid latitudes (deg)
    const double* col_lon;       // nij point longitudes
    const double* col_lat;       // nij point latitudes
    const double* col_dlon_zero; // nij  dlat_zero / cos(lat)  (speedy/letkf/letkf_obs.f90:86-88)
    double dlat_zero, dist_zero, dist_zerov, sigma_obs, sigma_obsv;
    const double* logpfm;        // nlev x nij  ln(ps_mean * sig)
    const double* var_local;     // 7 factors (u,v,t,q,rh,ps,rain)
    // binned obs (sorted order)
    const int* nobsgrd;          // nlon x nlat cumulative
    const int* o_typ;
    const double *o_sinlat, *o_coslat, *o_lonr, *o_lnlev, *o_lndat;
    // outputs
    int* cand_cnt;               // nij
    long* cand_off;              // nij + 1
    long* loc_off;               // nlev * nij (relative to the list buffer of the launch)
    int* loc_cnt;                // nlev * nij
    int* loc_idx;                // nlev * sum(cand_cnt) over the launch's columns
    double* loc_rho;
    int col_begin;               // first column of a fill launch
    long list_base;              // cand_off[col_begin] * nlev
};

struct L96CullArgs {
    int nx, ny;
    const int* gidx;             // ny grid index of each obs
    double xlocal, xmaxloc;
    long* loc_off;
    int* loc_cnt;
    int* loc_idx;                // nx * ny
    double* loc_rho;
};

struct ObsPrepArgs {
    int nobs, ne, nlon, nlat;
    const double* lon;
    const double* lat;
    const double *elm, *olon, *olat, *olev, *dat, *err;
    const int* qc;               // nullable (1 = good)
    const double* hxf;           // H(x) of member i for obs n at hxf[n*h_so + i*h_sm]
    long h_so, h_sm;
    double gross_error;
    double* tmp_mean;
    double* tmp_dep;
    unsigned long long* keys;    // padded to a power of two
    int* cell_cnt;               // nlon * nlat (zeroed before)
    // sorted outputs
    int* perm;
    double* o_hdxf;
    float* o_hdxf32;             // nullable
    double *o_dep, *o_rdiag, *o_err;
    int* o_typ;
    double *o_sinlat, *o_coslat, *o_lonr, *o_lnlev, *o_lndat;
    int* nobsgrd;
};

hipError_t lekf_cull_global_launch(const GlobalCullArgs& A, hipStream_t st);
hipError_t lekf_cull_global_count_launch(const GlobalCullArgs& A, hipStream_t st);
hipError_t lekf_cull_global_fill_launch(const GlobalCullArgs& A, int ncols, hipStream_t st);
hipError_t lekf_cull_l96_launch(const L96CullArgs& A, hipStream_t st);
hipError_t lekf_obs_prep_launch(const ObsPrepArgs& A, hipStream_t st);
hipError_t lekf_sort_keys_launch(unsigned long long* keys, long n2, hipStream_t st);
hipError_t lekf_obs_scatter_launch(const ObsPrepArgs& A, int nbinned, int ncell, hipStream_t st);
hipError_t lekf_logpfm_launch(const double* x2b, int nij, int nv2d, int ips, int ne,
                              const double* sig, int nlev, double* logpfm, hipStream_t st);

// solve (lekf_solve.hip)
size_t lekf_solve_lds_bytes(int k, int nvmax);
size_t lekf_solve_ws_bytes(long npt, int kmax);
int lekf_solve_nt(int k);
int lekf_solve_class(int k);
hipError_t lekf_solve_launch(const LekfJob* djobs, int njobs, long npt, int kmax, int nvmax,
                             int mode, void* ws, size_t ws_bytes, hipStream_t st);
