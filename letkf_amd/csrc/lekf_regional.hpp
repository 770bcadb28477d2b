// lekf_regional.hpp -- argument blocks of the regional obs-index and selection kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "lekf_types.h"

#define LEKF_REG_NB 4      // buckets (the reference's (element, obs type) pairs)
#define LEKF_REG_SEG 512   // rectangle rows per point
#define LEKF_CAP_LDS 1024  // caps up to this size keep their kept set in LDS

struct RegObsArgs {
    int nobs, ne, nx, ny;
    const double *ri, *rj, *z, *err, *dep;  // ri, rj: 1-based grid units; z: metres
    const int* bkt;
    const int* qc;                          // nullable: keep the obs whose qc is qc_keep
    int qc_keep;
    const double* hdxf;                     // obs perturbation of member i at hdxf[n*h_so + i*h_sm]
    long h_so, h_sm;
    unsigned long long* keys;               // padded to a power of two
    int* cell_cnt;                          // nx * ny, zeroed before
    // sorted outputs
    int* perm;
    double *o_ri, *o_rj, *o_z, *o_err, *o_rdiag, *o_dep, *o_hdxf;
    float* o_hdxf32;                        // nullable
    int* o_bkt;
};

struct RegionalArgs {
    int nx, ny, nlev, nbkt, nobs_max;
    double dx, dy, sigma_h, sigma_v, dlon_zero, dlat_zero;
    const double* zlev;                     // nlev level heights (m)
    const int* cellcum;                     // nx * ny inclusive cumulative obs counts, (j, i) order
    const double *ori, *orj, *oz, *oerr;    // sorted obs
    const int* obkt;
    // the sorted obs again, bucket by bucket (each bucket's obs in the sorted
    // order): bcc[t * nx * ny + c] = inclusive count of bucket t's obs through
    // cell c; bucket t starts at the sum of the earlier buckets' totals;
    // brec[4 p ..] = (ri, rj, z, err) of the obs at position p (32-byte
    // aligned); bo = the obs's position in the sorted order
    const int* bcc;
    const double* brec;
    const int* bo;
    // per bucket, the smallest err^2 of its obs (double bits)
    const unsigned long long* err2_min;
    // the fused selection's slot -> obs map, LEKF_CAP_LDS ints per wave, for
    // wc_so_waves waves (the launch holds no more waves than that)
    int* wc_so;
    long wc_so_waves;
    long pt_begin;                          // first global point of this launch
    const long* list_off;                   // per point list offsets (global numbering)
    long list_base;                         // list_off of the launch's first point
    long* loc_off;                          // per point, relative to the launch's list buffer
    int* loc_cnt;
    int* loc_idx;
    double* loc_rho;
    // per-workgroup scratch of the capped selection: one stream per bucket
    // (slot_cap entries each) of the point's accepted candidates in scan
    // order (key = rdiag bits, obs, rho)
    unsigned long long* s_key;
    int* s_ob;
    double* s_rho;
    long slot_cap;                          // entries per bucket stream
    // caps above LEKF_CAP_LDS: the kept set, 2 * nobs_max entries per
    // workgroup slot (null otherwise)
    unsigned long long* g_key;
    int* g_pos;
    int* g_slot;
};

// Nearest-first cap of per-point local lists (the global grid's nobs_max):
// keeps the cap entries of largest rho (ties: the earlier list position) of
// each input list, in the reference cap's storage-slot order, at
// out_*[b * cap].
struct ListCapArgs {
    int nij, nlev;                          // points b = lev * nij + ij
    int col_begin, ncols;                   // columns [col_begin, col_begin + ncols), all levels
    int cap;
    const long* in_off;
    const int* in_cnt;
    const int* in_idx;
    const double* in_rho;
    long* out_off;
    int* out_cnt;
    int* out_idx;
    double* out_rho;
    // caps above LEKF_CAP_LDS: the kept set, 2 * cap entries per workgroup
    // slot (null otherwise)
    unsigned long long* g_key;
    int* g_pos;
    int* g_slot;
};

hipError_t lekf_list_cap_launch(const ListCapArgs& A, int nslots, hipStream_t st);

hipError_t lekf_reg_keys_launch(const RegObsArgs& A, hipStream_t st);
// the bucket-by-bucket copy of the sorted obs (RegionalArgs bcc, bri .. bo):
// bcnt (nbkt * ncell ints) is work space
hipError_t lekf_reg_buckets_launch(const RegObsArgs& A, int nbinned, const int* cellcum, int nbkt,
                                   int* bcnt, int* bcc, double* brec, int* bo,
                                   unsigned long long* err2_min, hipStream_t st);
// the capped selection keeps its kept sets in LDS, one wave per bucket
inline bool lekf_reg_select_fused(int nobs_max) { return nobs_max > 0 && nobs_max <= LEKF_CAP_LDS; }
hipError_t lekf_reg_scatter_launch(const RegObsArgs& A, int nbinned, int ncell, int* cellcum,
                                   hipStream_t st);
// per point list capacity (rectangle count, capped) and its exclusive scan;
// maxcand <- the largest uncapped rectangle count (zeroed before)
hipError_t lekf_reg_capacity_launch(const RegionalArgs& A, int* capn, long* cap_off, int* maxcand,
                                    hipStream_t st);
// persistent selection over npt points with nslots workgroups (scratch slots)
hipError_t lekf_reg_select_launch(const RegionalArgs& A, long npt, int nslots, hipStream_t st);
int lekf_reg_select_slots();      // tuned count
int lekf_reg_select_slots_max();  // scratch is sized for this many
// waves the fused selection's slot maps are sized for
long lekf_reg_wave_slots_max();
