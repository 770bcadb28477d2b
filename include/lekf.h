/*
 * lekf.h -- C ABI of the MI355X LETKF local-analysis library (liblekf_hip.so).
 *
 * Drop-in boundary for the reference's LETKF hot path:
 *   letkf_core  common/common_letkf.f90:50-243   -> lekf_core_f64 (per point)
 *                                                  lekf_core_batch (many points, device memory)
 *   obs_local   speedy/letkf/letkf_tools.f90:272-603, lorenz96/letkf/letkf.f90:216-254,
 *               scale/letkf/letkf_tools.f90:853-1178 -> the selection inside lekf_das_run and
 *                                                  lekf_obs_local_* (device memory)
 *   set_letkf_obs speedy/letkf/letkf_obs.f90:48-335 -> lekf_set_obs_global / lekf_set_obs_l96
 *   das_letkf   speedy/letkf/letkf_tools.f90:45-267  -> lekf_*_create + lekf_set_ensemble +
 *                                                  lekf_das_run + lekf_get_analysis
 *
 * Plain pointers and sizes only.  Host arrays use the reference's Fortran
 * column-major layouts.  Every entry point returns a status code; the
 * Fortran shims map nonzero codes to the reference's STOP codes.
 * lekf_core_f64 is re-entrant and thread-safe (per-thread stream and
 * buffers), as OpenMP callers of letkf_core require
 * (scale/letkf/letkf_tools.f90:175).  A handle is used by one caller at a time.
 */
#ifndef LEKF_H
#define LEKF_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define LEKF_OK 0
#define LEKF_ERR_EIGEN 2      /* eigensolver failure: common/common_mtx.f90:61-64,75-78 STOP 2 */
#define LEKF_WARN_RANKDEF 4   /* lambda_min < lambda_max*sqrt(eps): the reference truncates the
                                 rank (common_mtx.f90:66-74) and produces NaN; we return finite
                                 values and flag the point */
#define LEKF_ERR_SIZE 99      /* inconsistent sizes: speedy/letkf/letkf_tools.f90:557-561 STOP 99 */
#define LEKF_ERR_HIP 100      /* HIP runtime error */
#define LEKF_ERR_ARG 101      /* invalid argument */
#define LEKF_ERR_NODEV 102    /* no GPU device */

/* arithmetic forms */
#define LEKF_PREC_F64 0     /* plain: all products FP64 */
#define LEKF_PREC_F32 1     /* mixed: Y^T R^-1 Y on FP32 MFMA, eigen / W / apply FP64 */
#define LEKF_PREC_BF16X3 2  /* mixed: Y^T R^-1 Y as split-BF16x3 on MFMA, FP32 accumulation */

/* inflation modes of the das entry points */
#define LEKF_INFL_ADAPTIVE 0  /* letkf_core's adaptive update kept (speedy/letkf/letkf_tools.f90:160-162) */
#define LEKF_INFL_FIXED 1     /* reset to infl_value after each point (lorenz96/letkf/letkf.f90:256) */

/* relaxation of the analysis (lekf_set_relax) */
#define LEKF_RELAX_NONE 0
#define LEKF_RELAX_RTPP 1  /* relaxation to prior perturbations (Zhang et al. 2005) */
#define LEKF_RELAX_RTPS 2  /* relaxation to prior spread (Whitaker and Hamill 2012) */

/* job flags (lekf_job.flags) */
#define LEKF_F_WLOC 1        /* rdiag already err^2/rho (rdiag_wloc=.true.) */
#define LEKF_F_MINFL 2       /* minfl present */
#define LEKF_F_FIXED_INFL 4  /* write infl_fixed as the updated inflation */
#define LEKF_F_TRANSM 8      /* return Wa in trans and w-bar in transm (transm present) */
#define LEKF_F_APPLY 16      /* apply Xa = xmean + Xb' T to the point's state rows */
#define LEKF_F_MIXED 32
#define LEKF_F_BF16X3 64

/* per-point status bits (lekf_job.status) */
#define LEKF_PT_RANKDEF 1
#define LEKF_PT_NOCONV 2
#define LEKF_PT_FP64 4       /* mixed forms: C was worse conditioned than LEKF_TUNE_MIXED_KAPPA
                                and the point was solved again from an FP64 Gram */

#define LEKF_NVMAX 16        /* state rows per point in the fused apply */
#define LEKF_KMAX 256        /* largest ensemble size */
#define LEKF_KSPLIT 95       /* largest ensemble size of the LDS/register eigensolver */

/*
 * One batched analysis job (device pointers).  The library fills pt_begin.
 * Y(o, m) = hdxf[o*y_so + m*y_sm]; point b uses local entries
 * loc_off[b] .. loc_off[b] + cnt (cnt = loc_cnt[b], or loc_off[b+1]-loc_off[b]
 * when loc_cnt is NULL), each an obs index and a localisation weight.
 * Fused apply rows: point b = lev*nij + ij owns nv3d rows at
 * x3b + (b*nv3d + v)*k and, at lev 0, nv2d rows at x2b + (ij*nv2d + v)*k.
 */
typedef struct lekf_job {
    long pt_begin;
    long pt0;                /* global index of the job's first point: per-point arrays
                                (loc_off, loc_cnt, infl, outputs, state rows) are
                                indexed by pt0 + local point */
    int npt;
    int k;
    int flags;
    int nij;
    int nv3d, nv2d;
    double minfl, infl_fixed;
    const double *hdxf;
    long y_so, y_sm;
    const float *hdxf32;     /* mixed forms: Y(o, m) = hdxf32[o*k + m] */
    const double *rdiag;
    const double *dep;
    const long *loc_off;
    const int *loc_cnt;      /* nullable */
    const int *loc_idx;
    const double *loc_rho;
    const double *loc_rdiag; /* nullable */
    const double *infl_in;   /* nullable: infl_fixed used */
    double *infl_out;        /* nullable */
    double *trans;           /* nullable, k x k column-major per point */
    double *transm;          /* nullable */
    double *pao;             /* nullable */
    int *status;             /* nullable */
    const double *x3b;
    double *x3a;
    const double *x2b;
    double *x2a;
    const uint32_t *upd_mask; /* nullable: bit v of upd_mask[lev] set = 3-D variable v updated */
} lekf_job;

/* --- 1. per-point drop-in --------------------------------------------------
 * letkf_core(ne,nobs,nobsl,hdxb,rdiag,rloc,dep,parm_infl,trans,transm,pao,
 *            rdiag_wloc,minfl)   common/common_letkf.f90:50
 * hdxb is nobs x ne column-major (first nobsl rows used).  transm, pao and
 * minfl may be NULL (absent OPTIONAL arguments).  Host memory. */
int lekf_core_f64(int ne, int nobs, int nobsl, const double *hdxb, const double *rdiag,
                  const double *rloc, const double *dep, double *parm_infl, double *trans,
                  double *transm, double *pao, int rdiag_wloc, const double *minfl);

/* --- 2. batched local solve ----------------------------------------------
 * Solves every point of njobs jobs (device memory) on `stream` (hipStream_t,
 * NULL = default stream); prec selects the Gram arithmetic.  Asynchronous. */
int lekf_core_batch(const lekf_job *jobs, int njobs, int prec, void *stream);

/* --- 2b. batched local solve without allocation -----------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same solve as lekf_core_batch, bit for bit, with
 * the job descriptors and the Gram workspace in a fixed device array of
 * LEKF_NOALLOC_BYTES per GPU (held until the process ends) instead of
 * library allocations; it calls no hipMalloc/hipFree.  At most
 * LEKF_NOALLOC_MAX_JOBS jobs; a point record (k > 32) larger than the
 * workspace part returns LEKF_ERR_SIZE (lekf_last_error says which).  Calls
 * that overlap on one GPU share the array and race: serialise them. */
#define LEKF_NOALLOC_BYTES (256L << 20)
#define LEKF_NOALLOC_MAX_JOBS 1024
int lekf_core_batch_noalloc(const lekf_job *jobs, int njobs, int prec, void *stream);

/* --- 3. selection on device memory ---------------------------------------
 * Lorenz-96 rule (lorenz96/letkf/letkf.f90:216-254): obs i of grid index
 * gidx[i]; point ix keeps |periodic distance| < 2*sqrt(10/3)*xlocal and
 * rho >= 1e-4.  Lists at loc_off[ix] = ix*ny. */
int lekf_obs_local_l96(int nx, int ny, const int *gidx, double xlocal, long *loc_off, int *loc_cnt,
                       int *loc_idx, double *loc_rho, void *stream);

/* --- 4. das_letkf-level handle -------------------------------------------- */
typedef struct lekf_handle lekf_handle;

typedef struct lekf_global_cfg {
    int nlon, nlat, nlev;     /* grid; ij = i + nlon*j */
    const double *lon;        /* nlon longitudes (deg), ascending in [0, 360) */
    const double *lat;        /* nlat latitudes (deg), ascending */
    const double *sig;        /* nlev sigma levels (p_full = ps*sig) */
    int ne;                   /* ensemble size */
    int nv3d, nv2d;           /* 3-D variables, 2-D variables (updated with the level-1 W) */
    int ips;                  /* surface pressure among the 2-D variables */
    int iq;                   /* specific humidity among the 3-D variables, -1: none */
    int q_top_lev;            /* 1-based level from which q keeps the background; 0: off */
    double sigma_obs;         /* horizontal localisation scale (m) */
    double sigma_obsv;        /* vertical localisation scale (ln p) */
    double gross_error;       /* QC: |d| > gross_error * err is rejected */
    double var_local[7];      /* localisation factor per obs type u,v,t,q,rh,ps,rain */
    int prec;                 /* LEKF_PREC_* */
    int infl_mode;            /* LEKF_INFL_* */
    double infl_value;        /* fixed value, or the initial adaptive field value */
    int nobs_max;             /* > 0: keep at most nobs_max local obs per point, nearest
                               * first (largest rho; ties: the earlier obs in scan order),
                               * by the reference cap's insertion sort and storage-slot
                               * order (scale/letkf/letkf_tools.f90:1041-1156, key -rho);
                               * 0: no cap */
} lekf_global_cfg;

typedef struct lekf_l96_cfg {
    int nx;                   /* grid points */
    int ne;
    double xlocal;            /* localisation scale (grid points) */
    int prec;
    int infl_mode;
    double infl_value;
} lekf_l96_cfg;

/* Regional cartesian grid (SCALE-style, one subdomain): point (i, j, lev) at
 * grid units ri = i, rj = j (1-based) and height zlev[lev]; radar-type
 * vertical distance |z_o - z| (scale/letkf/letkf_tools.f90:992-1000). */
typedef struct lekf_regional_cfg {
    int nx, ny, nlev;
    double dx, dy;            /* grid spacing (m) */
    const double *zlev;       /* nlev level heights (m) */
    int ne;
    int nv3d;                 /* state variables (all 3-D) */
    double sigma_h, sigma_v;  /* localisation scales (m) */
    int nobs_max;             /* MAX_NOBS_PER_GRID per bucket (kept obs emitted bucket by
                               * bucket in the reference's storage-slot order,
                               * scale/letkf/letkf_tools.f90:1041-1156); <= 0: no cap */
    int prec;
    int infl_mode;
    double infl_value;
    double minfl;             /* INFL_MUL_MIN; <= 0: off */
    long list_budget;         /* local-list entries held at once (0: library default) */
} lekf_regional_cfg;

lekf_handle *lekf_global_create(const lekf_global_cfg *cfg, int *status);
lekf_handle *lekf_regional_create(const lekf_regional_cfg *cfg, int *status);
lekf_handle *lekf_l96_create(const lekf_l96_cfg *cfg, int *status);
void lekf_destroy(lekf_handle *h);

/* background ensemble, reference layout: gues3d(nij,nlev,ne,nv3d),
 * gues2d(nij,ne,nv2d), column-major (L96: nij = nx, nlev = 1, nv3d = 1). */
int lekf_set_ensemble(lekf_handle *h, const double *gues3d, const double *gues2d);
/* inflation field infl3d(nij,nlev) (and infl2d(nij,nv2d), may be NULL) */
int lekf_set_infl(lekf_handle *h, const double *infl3d, const double *infl2d);
/* set_letkf_obs (speedy/letkf/letkf_obs.f90:48-335): obs records and H(x) of
 * every member, hxf(nobs,ne) column-major; qc may be NULL (all good). */
int lekf_set_obs_global(lekf_handle *h, int nobs, const double *elm, const double *lon,
                        const double *lat, const double *lev, const double *dat,
                        const double *err, const int *qc, const double *hxf);
/* The same with H(x) computed on the GPU from the resident background by the
 * reference obs operator (obsope + Trans_XtoY, speedy/letkf/obsope.f90:35-111,
 * speedy/common/common_obs_speedy.f90:30-282): u, v, t, q (3-D variables
 * 0, 1, 2, iq) and ps (2-D variable ips); other elements are rejected. */
int lekf_set_obs_global_hx(lekf_handle *h, int nobs, const double *elm, const double *lon,
                           const double *lat, const double *lev, const double *dat,
                           const double *err);
/* The obs operator alone (obsope, speedy/letkf/obsope.f90:54-105): H(x) of
 * every member hxf(nobs,ne) column-major (0 where rejected), qc(nobs) = 1 when
 * every member accepts the obs, and optionally the grid coordinates ri, rj
 * (nobs) and rk(nobs,ne) of phys2ijk (common_obs_speedy.f90:137-211). */
int lekf_obsope_global(lekf_handle *h, int nobs, const double *elm, const double *lon,
                       const double *lat, const double *lev, const double *dat, double *hxf,
                       int *qc, double *ri, double *rj, double *rk);
/* SCALE's radar obs operator on a regional handle's resident ensemble.
 * The 3-D variables var[0..10] of the handle hold u, v, w, T, p, q, qc, qr, qi,
 * qs, qg (SCALE's iv3dd order, scale/common/common_scale.f90:103-113).
 * method: METHOD_REF_CALC 1, 2 or 3; use_terminal_velocity, min_ref_dbz,
 * low_ref_shift and zmax: USE_TERMINAL_VELOCITY, MIN_RADAR_REF_DBZ,
 * LOW_REF_SHIFT and RADAR_ZMAX (scale/common/common_nml.f90:173-184; defaults
 * 3, 0, 0.0, 0.0, 99e3). */
typedef struct lekf_radar_cfg {
    int var[11];
    int method;
    int use_terminal_velocity;
    double min_ref_dbz, low_ref_shift, zmax;
} lekf_radar_cfg;
/* heights (m) of the model levels, hgt(nij, nlev) column-major, ascending with
 * the level and the same for every member (SCALE's z_full, iv3dd_hgt) --
 * regional handles of at least two levels only */
int lekf_set_heights(lekf_handle *h, const double *hgt);
/* H(x) of radar reflectivity (elm 4001, dBZ) and Doppler velocity (4002, m/s)
 * for every member: the checks of SCALE's obsope (RADAR_ZMAX, phys2ijkz,
 * scale/obs/obsope_tools.f90:353-380) and Trans_XtoY_radar / calc_ref_vr
 * (scale/common/common_obs_scale.f90:338-476,609-960).  Per obs: element,
 * lon, lat (deg), height z (m), 1-based grid coordinates ri, rj (the map
 * projection's, 1 < ri <= nx, 1 < rj <= ny), the radar's lon, lat and height.
 * Out: hxf(nobs,ne) and qc(nobs,ne) column-major (qc 0 good; 19 above zmax,
 * 20 above the top, 21 below the lowest valid level, 22 outside the grid or at
 * the radar, 90 not a radar element), rk(nobs) (may be NULL). */
int lekf_obsope_radar(lekf_handle *h, const lekf_radar_cfg *cfg, int nobs, const double *elm,
                      const double *lon, const double *lat, const double *z, const double *ri,
                      const double *rj, const double *radar_lon, const double *radar_lat,
                      const double *radar_z, double *hxf, int *qc, double *rk);
/* Options of SCALE's departure and QC step: the PARAM_LETKF_OBS and
 * PARAM_LETKF_RADAR namelist values of scale/common/common_nml.f90:83-91,
 * 140-146,165-174 (defaults 1, 1, 0, 0, 1, 1, 15.0, 0.0, 0.0, 5.0, 3.0, 5.0 and
 * -1.0 for every element's gross-error factor; a factor below zero stands for
 * gross_error, common_nml.f90:348-371). */
typedef struct lekf_scale_qc_cfg {
    int use_radar_ref, use_radar_vr;
    int use_obserr_radar_ref, use_obserr_radar_vr;
    int min_radar_ref_member, min_radar_ref_member_obsref;
    double radar_ref_thres_dbz, min_radar_ref_dbz, low_ref_shift;
    double obserr_radar_ref, obserr_radar_vr;
    double gross_error, gross_error_rain, gross_error_radar_ref, gross_error_radar_vr,
        gross_error_radar_prh, gross_error_tcx, gross_error_tcy, gross_error_tcp;
} lekf_scale_qc_cfg;
/* SCALE's departure and QC step of set_letkf_obs on the GPU
 * (scale/letkf/letkf_obs.f90:320-343,357,476-508,560-645; replaces the loop
 * between the obs preprocessing and the sorting).  Per obs: element, value as
 * observed (reflectivity in mm^6 m^-3; dat_out gives it in dBZ, undef -9.99e33
 * outside [0, 1e10)), error (err_out: replaced by OBSERR_RADAR_REF / _VR when
 * selected); qc_member(nobs,ne) the operator's qc of each member (NULL: all
 * good), merged by the maximum; hx(nobs,ne) H(x) of each member, replaced by
 * the perturbations H(x) - mean where the obs's qc comes out good (rows of
 * other obs are left as they were).  dep = y - mean H(x) (0 unless good), qc
 * 0 good, 5 gross error, 12 too few members with echo, 50 bad obs, 90 element
 * not used, else the operator's.  Column-major arrays, ne = the handle's
 * members.  Himawari-8 obs (elm 8800) are rejected with LEKF_ERR_ARG: their
 * QC needs the RTTOV operator's cloud flags. */
int lekf_obs_departure_scale(lekf_handle *h, const lekf_scale_qc_cfg *cfg, int nobs,
                             const double *elm, const double *dat, const double *err,
                             const int *qc_member, double *hx, double *dat_out, double *err_out,
                             double *dep, int *qc);
/* Radar obs of a regional handle set up without leaving the device: the
 * radar operator (as lekf_obsope_radar), the departure and QC step (as
 * lekf_obs_departure_scale, with the members' qc and H(x) of the operator)
 * and the regional obs index (as lekf_set_obs_regional), in place of SCALE's
 * obsope + set_letkf_obs for radar files (scale/obs/obsope_tools.f90:353-380,
 * scale/letkf/letkf_obs.f90:320-830).  Per obs: the operator's inputs, the
 * value as observed (reflectivity in mm^6 m^-3), its error and bucket (the
 * (element, obs type) index of the cap, 0..3).  qc_out(nobs) (may be NULL):
 * the qc of each obs. */
int lekf_set_obs_radar(lekf_handle *h, const lekf_radar_cfg *rcfg, const lekf_scale_qc_cfg *qcfg,
                       int nobs, const double *elm, const double *lon, const double *lat,
                       const double *z, const double *ri, const double *rj,
                       const double *radar_lon, const double *radar_lat, const double *radar_z,
                       const double *dat, const double *err, const int *bucket, int *qc_out);
/* Regional obs after the obs operator (SCALE's obsda): grid-unit position
 * (ri, rj), height z (m), error, departure dep = y - mean H(x), bucket
 * (the (element, obs type) index, 0..3), qc (1: use the obs, any other value:
 * skip it, as the obs files' flag; NULL: use all; SCALE's qc codes, where 0 is
 * good, map to qc == 0 ? 1 : 0) and the obs-space
 * perturbations hdxf(nobs,ne) column-major (scale/letkf/letkf_obs.f90:733-830). */
int lekf_set_obs_regional(lekf_handle *h, int nobs, const double *ri, const double *rj,
                          const double *z, const double *err, const double *dep, const int *bucket,
                          const int *qc, const double *hdxf);
/* Lorenz-96 obs: grid index gidx (0-based), value y, error err, hxf(ny,ne)
 * column-major.  hxf NULL: H(x) is taken on the GPU from the resident
 * background (H selects grid points, lorenz96/obs/h_ope.f90:14-39). */
int lekf_set_obs_l96(lekf_handle *h, int ny, const int *gidx, const double *y, const double *err,
                     const double *hxf);
/* Lorenz-96 forecast of a served cycle: kt RK4 steps (tinteg_rk4 /
 * lorenz96_core, lorenz96/model/lorenz96.f90:93-128,290-309) of every member
 * of the latest analysis, with time step dt and forcing force; the forecast
 * becomes the background of the next cycle (the analysis stays readable) and
 * the updated inflation its input.  Each member's trajectory is bit-identical
 * to the reference model's. */
int lekf_l96_forecast(lekf_handle *h, int kt, double dt, double force);
/* one analysis (selection + local solves + apply) on the handle's stream;
 * asynchronous.  The analysis stays on the device. */
int lekf_das_run(lekf_handle *h);
/* enable != 0: lekf_das_run replays the cycle's kernel launches as one HIP
 * graph.  The first run after enabling or after any lekf_set_* /
 * lekf_cycle_swap runs eagerly and captures the graph; later runs launch it.
 * The graph holds no memset or allocation. */
int lekf_set_graph(lekf_handle *h, int enable);
/* Driver options of the SCALE/GFS das_letkf applied after the local solves
 * (scale/letkf/letkf_tools.f90:215-268): mode LEKF_RELAX_RTPP relaxes the
 * analysis perturbations to the background's with weight alpha
 * (RELAX_ALPHA, weight_RTPP :1224-1236), LEKF_RELAX_RTPS relaxes the spread
 * (RELAX_ALPHA_SPREAD, weight_RTPS :1240-1268); q_sprd_max > 0 limits
 * spread / mean of 3-D row iq (Q_SPRD_MAX :255-268); boundary_width > 0
 * (regional handles, metres) scales the update down within that distance of
 * the lateral boundary (relax_beta's BOUNDARY_BUFFER_WIDTH :1203-1216).
 * Default: all off. */
int lekf_set_relax(lekf_handle *h, int mode, double alpha, double q_sprd_max, int iq,
                   double boundary_width);
/* the two halves of lekf_das_run: selection kernels, then the local solves + apply */
int lekf_das_select(lekf_handle *h);
int lekf_das_solve(lekf_handle *h);
/* run the handle's work on a caller stream (hipStream_t; NULL = default stream) */
int lekf_set_stream(lekf_handle *h, void *stream);
/* batch the local solves of several handles into shared launches */
int lekf_das_run_many(lekf_handle **hs, int n);
/* the same batching, enqueued on one caller stream (hipStream_t) without
 * waiting: every handle's selection and the shared solve launches go to
 * stream; the handles keep running on it afterwards (lekf_set_stream moves
 * them).  Completion: record an event on stream (lekf_event_record) and wait
 * or poll (lekf_event_query).  Calls that overlap must use distinct streams
 * and distinct handles. */
int lekf_das_run_batch(lekf_handle **hs, int n, void *stream);
int lekf_sync(lekf_handle *h);
int lekf_get_analysis(lekf_handle *h, double *anal3d, double *anal2d);
int lekf_get_infl(lekf_handle *h, double *infl3d, double *infl2d);
int lekf_get_nobsl(lekf_handle *h, int *nobsl);
/* local list of point b (= lev*nij + ij): up to cap entries of sorted-obs index,
 * rho; returns the count (or a negative status).  A regional handle whose lists
 * were built in several chunks selects the chunk of b again first. */
int lekf_get_local(lekf_handle *h, long b, int cap, int *idx, double *rho);
/* sorted-obs permutation (input index of each binned obs) and the cumulative
 * obs count through each cell: nobsgrd(nlon,nlat) for the global grid,
 * (nx,ny) for the regional grid */
int lekf_get_obs_index(lekf_handle *h, int *perm, int *nobsgrd, int *nbinned);
/* number of points flagged rank-deficient / not converged in the last run */
int lekf_get_flags(lekf_handle *h, long *n_rankdef, long *n_noconv);
/* number of points of the last run solved again in FP64 (LEKF_PT_FP64) */
int lekf_get_fp64_count(lekf_handle *h, long *n_fp64);
/* use the analysis as the next background (cycling without host traffic) */
int lekf_cycle_swap(lekf_handle *h);
/* device pointers of the resident ensembles (for device-side pipelines) */
int lekf_device_state(lekf_handle *h, double **x3b, double **x3a, double **x2b, double **x2a);
/* bytes of device memory the handle holds */
long lekf_device_bytes(lekf_handle *h);
/* hipStream_t of the handle */
void *lekf_stream(lekf_handle *h);

/* --- 5. device memory, streams and events ----------------------------------
 * For hosts without a GPU runtime of their own (Fortran, Python/numpy):
 * the batched entry points take device pointers from here.  No counterpart
 * in the reference, which has no device.  kind: 0 host->device, 1
 * device->host (both synchronous), 2 device->device (asynchronous). */
void *lekf_dev_malloc(size_t bytes);
int lekf_dev_free(void *p);
int lekf_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream);
int lekf_memset(void *dst, int value, size_t bytes, void *stream);
void *lekf_stream_create(void);
int lekf_stream_destroy(void *stream);
int lekf_stream_sync(void *stream);
void *lekf_event_create(void);
int lekf_event_destroy(void *ev);
int lekf_event_record(void *ev, void *stream);
int lekf_event_elapsed_ms(void *ev0, void *ev1, float *ms);
/* 1 if all work before the event has completed, 0 if not, negative on error */
int lekf_event_query(void *ev);
int lekf_event_sync(void *ev);

/* launch tuning (process-wide; read when a launch is issued).  The values
 * change kernel shapes only: the workspace size and the selection slot count
 * leave every result bit-identical, the Gram wave count changes the order of
 * the partial-sum reduction only, the eigensolver choice agrees to rounding.  Handles recapture their graphs after a
 * change.  lekf_tune_get returns the value, or a negative error code. */
#define LEKF_TUNE_GRAM_WAVES 0 /* waves per point of the k > 32 Gram kernel: 2, 4 or 8 */
#define LEKF_TUNE_GRAM_WS_MB 1 /* Gram workspace per launch, MiB, 16 .. 65536 (16384) */
#define LEKF_TUNE_SEL_PER_CU 2 /* capped-selection workgroups per CU: 1 .. 4 */
#define LEKF_TUNE_EIG_SPLIT 3  /* 32 < k <= 64: 1 = tridiagonalisation, QL (one lane per point)
                                  and apply kernels (the apply pushes the state rows through
                                  Q and the recorded rotations instead of forming the
                                  eigenvectors, unless trans or pao is asked for), 0 = one
                                  fused eigen kernel; the two agree to rounding */
#define LEKF_TUNE_SPLIT_RCAP 4 /* rotations a point's QL stream may hold in the three-kernel
                                  eigensolver; a point whose stream is longer is solved by
                                  the fused kernel's steps instead; 0: the default 2k^2 + 8k
                                  (smaller values test that fallback) */
#define LEKF_TUNE_MIXED_KAPPA 5 /* split-BF16x3 and FP32 forms on the three-kernel eigensolver
                                   (32 < k <= 64): a point whose C + (k-1)/alpha I has condition
                                   number above this value (16 times it for FP32) is solved
                                   again from an FP64 Gram (status bit LEKF_PT_FP64);
                                   0 = never (0 .. 1e12; 500) */
int lekf_tune_set(int param, long value);
long lekf_tune_get(int param);

/* per-kernel timing, for measurement (no counterpart in the reference, which
 * times whole phases with its TIMER prints, speedy/letkf/letkf.f90:31,70-72).
 * lekf_ktime_enable(1) clears the record and starts recording a HIP event
 * pair on the launch's stream around each launch of the named kernels
 * ("select" the regional selection, "list_cap" the global cap, "cull" the
 * global candidate scan, "solve" the fused k <= 32 kernel, "gram", "eig",
 * "eig_tri", "eig_ql", "eig_apply", "big"); launches captured into a graph
 * are not recorded.  lekf_ktime_read waits for the recorded launches of one
 * name and returns their summed time and count.  Process-wide; 0 stops. */
int lekf_ktime_enable(int on);
int lekf_ktime_read(const char *name, double *total_ms, long *launches);

/* library info */
int lekf_device_count(void);
/* free and total device memory of the current GPU, bytes */
int lekf_mem_info(size_t *free_bytes, size_t *total_bytes);
/* device bytes of the local-solve workspace a grid of npt points at ensemble
 * size k takes under the current tuning (0 for k <= 32), for callers that
 * budget device memory (no counterpart in the reference, which has no device
 * workspace) */
long lekf_solve_workspace_bytes(long npt, int k);
const char *lekf_version(void);
const char *lekf_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
