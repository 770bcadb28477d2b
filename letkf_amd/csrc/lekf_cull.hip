// lekf_cull.hip -- local observation selection and the obs index build (gfx950).
//
// Global lat/lon grids (SPEEDY/GFS das_letkf semantics):
//   * obs index build (speedy/letkf/letkf_obs.f90:146-313): ensemble mean of
//     H(x), perturbations, departures, gross-error QC, then a stable sort of
//     the surviving obs into lat band / lon cell order with the cumulative
//     per-cell index nobsgrd.  The sort is a bitonic sort of unique 64-bit
//     keys (cell << 32 | input index), which gives exactly the stable order.
//   * cull (speedy/letkf/letkf_tools.f90:272-603): one workgroup per grid
//     column enumerates the reference's lat/lon rectangle of bins (with the
//     0/360 wrap cases) in the reference's scan order, computes each
//     candidate's great-circle distance once (common/common.f90:401-424)
//     and tests every level of the column against it, compacting the kept
//     obs per level with wave ballots so the list order is the scan order.
// Lorenz-96 (lorenz96/letkf/letkf.f90:216-254): periodic grid distance.
#include "lekf_dev.hpp"
#include "lekf_cull.hpp"
#include "lekf_ktime.hpp"

// Every FP64 operation in this file is rounded as written: no FMA contraction
// (the reference's results are reproduced operation by operation).
#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ unsigned long long lane_mask_lt(int lane)
{
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// DO j=1,nmax: IF (x < tab(j+1)) EXIT  -> 1-based j in 1..nmax+1
__device__ __forceinline__ int first_lt(double x, const double* tab, int nmax)
{
    int j = 1;
    for (; j <= nmax; ++j)
        if (x < tab[j]) break;
    return j;
}

struct RectSegs {
    int nseg;
    int jmin, jmax;
    int ia[2], ib[2];  // lon ranges (1-based, inclusive)
};

// speedy/letkf/letkf_tools.f90:302-451
__device__ RectSegs rect_ranges(double lon1, double lat1, double dlon_zero, double dlat_zero,
                                const double* lon, const double* lat, int nlon, int nlat)
{
    RectSegs R;
    double minlon = lon1 - dlon_zero, maxlon = lon1 + dlon_zero;
    const double minlat = lat1 - dlat_zero, maxlat = lat1 + dlat_zero;
    if (maxlon - minlon >= 360.0) {
        minlon = 0.0;
        maxlon = 360.0;
    }
    R.jmin = first_lt(minlat, lat, nlat - 2);
    R.jmax = first_lt(maxlat, lat, nlat - 2);
    R.nseg = 1;
    int imin, imax;
    if (minlon >= 0 && maxlon <= 360.0) {
        imin = first_lt(minlon, lon, nlon - 1);
        imax = first_lt(maxlon, lon, nlon - 1);
        R.ia[0] = imin;
        R.ib[0] = imax;
    } else if (minlon >= 0 && maxlon > 360.0) {
        imin = first_lt(minlon, lon, nlon - 1);
        maxlon -= 360.0;
        if (maxlon > 360.0) {
            R.ia[0] = 1;
            R.ib[0] = nlon;
        } else {
            imax = first_lt(maxlon, lon, nlon - 1);
            if (imax > imin) {
                R.ia[0] = 1;
                R.ib[0] = nlon;
            } else {
                R.nseg = 2;
                R.ia[0] = 1;
                R.ib[0] = imax;
                R.ia[1] = imin;
                R.ib[1] = nlon;
            }
        }
    } else if (minlon < 0 && maxlon <= 360.0) {
        imax = first_lt(maxlon, lon, nlon - 1);
        minlon += 360.0;
        if (minlon < 0) {
            R.ia[0] = 1;
            R.ib[0] = nlon;
        } else {
            imin = first_lt(minlon, lon, nlon - 1);
            if (imin < imax) {
                R.ia[0] = 1;
                R.ib[0] = nlon;
            } else {
                R.nseg = 2;
                R.ia[0] = 1;
                R.ib[0] = imax;
                R.ia[1] = imin;
                R.ib[1] = nlon;
            }
        }
    } else {
        maxlon -= 360.0;
        minlon += 360.0;
        if (maxlon > 360.0 || minlon < 0) {
            R.ia[0] = 1;
            R.ib[0] = nlon;
        } else {
            imin = first_lt(minlon, lon, nlon - 1);
            imax = first_lt(maxlon, lon, nlon - 1);
            if (imin > imax) {
                R.ia[0] = 1;
                R.ib[0] = nlon;
            } else {
                R.ia[0] = imin;
                R.ib[0] = imax;
            }
        }
    }
    return R;
}

// [ib, ie) of band j (1-based) and lon range [i0, i1] (speedy/letkf/letkf_tools.f90:575-603)
__device__ __forceinline__ void seg_range(const int* nobsgrd, int nlon, int i0, int i1, int j,
                                          int& beg, int& end)
{
    if (i0 > 1)
        beg = nobsgrd[(i0 - 2) + (j - 1) * nlon];
    else
        beg = (j > 1) ? nobsgrd[(nlon - 1) + (j - 2) * nlon] : 0;
    end = nobsgrd[(i1 - 1) + (j - 1) * nlon];
    if (end < beg) end = beg;
}

__global__ void cull_global_count(GlobalCullArgs A)
{
    const int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= A.nij) return;
    const RectSegs R = rect_ranges(A.col_lon[ij], A.col_lat[ij], A.col_dlon_zero[ij], A.dlat_zero,
                                   A.lon, A.lat, A.nlon, A.nlat);
    long cnt = 0;
    for (int s = 0; s < R.nseg; ++s)
        for (int j = R.jmin; j <= R.jmax; ++j) {
            int b, e;
            seg_range(A.nobsgrd, A.nlon, R.ia[s], R.ib[s], j, b, e);
            cnt += e - b;
        }
    A.cand_cnt[ij] = (int)cnt;
}

// no fused multiply-add: the reference's x86 build evaluates these
// expressions with separate roundings
__device__ __forceinline__ double dist_gc(double sin_o, double cos_o, double lonr_o, double sin_g,
                                          double cos_g, double lonr_g)
{
    const double c = cos(rn_sub(lonr_g, lonr_o));
    double cosd = rn_add(rn_mul(sin_o, sin_g), rn_mul(rn_mul(cos_o, cos_g), c));
    cosd = cosd < 1.0 ? cosd : 1.0;
    cosd = cosd > -1.0 ? cosd : -1.0;
    return rn_mul(acos(cosd), LEKF_RE_REF);
}

__global__ __launch_bounds__(64) void cull_global_fill(GlobalCullArgs A)
{
    const int ij = A.col_begin + blockIdx.x;
    const int lane = threadIdx.x;
    __shared__ int s_cnt[LEKF_MAX_LEV];
    __shared__ int s_seg_beg[2 * LEKF_MAX_BANDS];
    __shared__ int s_seg_pre[2 * LEKF_MAX_BANDS + 1];
    const int nlev = A.nlev;
    const RectSegs R = rect_ranges(A.col_lon[ij], A.col_lat[ij], A.col_dlon_zero[ij], A.dlat_zero,
                                   A.lon, A.lat, A.nlon, A.nlat);
    const int nb = R.jmax - R.jmin + 1;
    const int nsegs = R.nseg * nb;
    if (lane == 0) {
        int acc = 0;
        for (int s = 0; s < R.nseg; ++s)
            for (int j = R.jmin; j <= R.jmax; ++j) {
                int b, e;
                seg_range(A.nobsgrd, A.nlon, R.ia[s], R.ib[s], j, b, e);
                const int t = s * nb + (j - R.jmin);
                s_seg_beg[t] = b;
                s_seg_pre[t] = acc;
                acc += e - b;
            }
        s_seg_pre[nsegs] = acc;
    }
    for (int l = lane; l < nlev; l += 64) s_cnt[l] = 0;
    __syncthreads();
    // this workgroup's levels: group blockIdx.y of gridDim.y equal runs
    const int lpg = (nlev + (int)gridDim.y - 1) / (int)gridDim.y;
    const int l_lo = (int)blockIdx.y * lpg;
    const int l_hi = l_lo + lpg < nlev ? l_lo + lpg : nlev;
    const int ncand = s_seg_pre[nsegs];
    const int cap = A.cand_cnt[ij];
    const long base = A.cand_off[ij] * (long)nlev - A.list_base;
    const double lat1 = A.col_lat[ij];
    const double r180 = 1.0 / 180.0;
    const double lonr_g = rn_mul(rn_mul(A.col_lon[ij], LEKF_PI_REF), r180);
    const double latr_g = rn_mul(rn_mul(lat1, LEKF_PI_REF), r180);
    const double sin_g = sin(latr_g), cos_g = cos(latr_g);
    int seg = 0;
    for (int q0 = 0; q0 < ncand; q0 += 64) {
        const int q = q0 + lane;
        const bool valid = q < ncand;
        int o = 0;
        double dist = 0.0;
        bool hok = false;
        if (valid) {
            while (s_seg_pre[seg + 1] <= q) ++seg;
            o = s_seg_beg[seg] + (q - s_seg_pre[seg]);
            dist = dist_gc(A.o_sinlat[o], A.o_coslat[o], A.o_lonr[o], sin_g, cos_g, lonr_g);
            hok = !(dist > A.dist_zero);
        }
        const int typ = valid ? A.o_typ[o] : 0;
        const double vl = (valid && typ >= 0) ? A.var_local[typ] : 0.0;
        hok = hok && (vl >= 2.2250738585072014e-308);
        for (int l = l_lo; l < l_hi; ++l) {
            bool sel = false;
            double dlev = 0.0;
            if (hok) {
                if (typ == LEKF_TYP_PS) {
                    dlev = (l > 0) ? fabs(rn_sub(A.o_lndat[o], A.logpfm[(size_t)l * A.nij + ij]))
                                   : 0.0;
                } else {
                    dlev = fabs(rn_sub(A.o_lnlev[o], A.logpfm[(size_t)l * A.nij + ij]));
                }
                sel = !(dlev > A.dist_zerov);
            }
            const unsigned long long m = __ballot(sel);
            if (m == 0ull) continue;
            const int c0 = s_cnt[l];
            if (sel) {
                const int pos = c0 + __popcll(m & lane_mask_lt(lane));
                const long w = base + (long)l * cap + pos;
                const double a = dist / A.sigma_obs, bb = dlev / A.sigma_obsv;
                A.loc_idx[w] = o;
                A.loc_rho[w] = exp(-0.5 * (a * a + bb * bb)) * vl;
            }
            __syncthreads();
            if (lane == 0) s_cnt[l] = c0 + __popcll(m);
            __syncthreads();
        }
    }
    __syncthreads();
    for (int l = l_lo + lane; l < l_hi; l += 64) {
        const long b = (long)l * A.nij + ij;
        A.loc_off[b] = base + (long)l * cap;
        A.loc_cnt[b] = s_cnt[l];
    }
}

// Lorenz-96: point ix keeps obs i with periodic |gidx_i - ix| < xmaxloc and
// rho >= 1e-4 (lorenz96/letkf/letkf.f90:226-251).  One wave per point.
__global__ __launch_bounds__(64) void cull_l96(L96CullArgs A)
{
    const int ix = blockIdx.x;
    const int lane = threadIdx.x;
    __shared__ int s_n;
    if (lane == 0) s_n = 0;
    __syncthreads();
    const long base = (long)ix * A.ny;
    for (int q0 = 0; q0 < A.ny; q0 += 64) {
        const int i = q0 + lane;
        bool sel = false;
        double r = 0.0;
        if (i < A.ny) {
            const int dj = abs(A.gidx[i] - ix);
            const int dm = (A.nx - dj) < dj ? (A.nx - dj) : dj;
            const double dist = (double)dm;
            if (dist < A.xmaxloc) {
                const double a = dist / A.xlocal;
                r = exp(-0.5 * (a * a)) * exp(-0.5 * 0.0);
                sel = !(r < 0.0001);
            }
        }
        const unsigned long long m = __ballot(sel);
        const int c0 = s_n;
        if (sel) {
            const long w = base + c0 + __popcll(m & lane_mask_lt(lane));
            A.loc_idx[w] = i;
            A.loc_rho[w] = r;
        }
        __syncthreads();
        if (lane == 0) s_n = c0 + __popcll(m);
        __syncthreads();
    }
    if (lane == 0) {
        A.loc_off[ix] = base;
        A.loc_cnt[ix] = s_n;
    }
}

// exclusive prefix sum of n ints into long out[0..n] (one workgroup)
__global__ __launch_bounds__(1024) void scan_ex(const int* in, long* out, int n)
{
    __shared__ long s[1024];
    long carry = 0;
    for (int base = 0; base < n; base += 1024) {
        const int i = base + threadIdx.x;
        const long v = i < n ? in[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const long t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < n) out[i] = carry + s[threadIdx.x] - v;
        const long tot = s[1023];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

// ---------------------------------------------------------------------------
// obs index build (speedy/letkf/letkf_obs.f90:157-174, 243-313)
// ---------------------------------------------------------------------------
__global__ void obs_prep(ObsPrepArgs A)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= A.nobs) return;
    unsigned long long key = ~0ull;
    // ensemble mean of H(x) in member order, perturbations, departure
    double m = A.hxf[(size_t)n * A.h_so];
    for (int i = 1; i < A.ne; ++i) m += A.hxf[(size_t)n * A.h_so + (size_t)i * A.h_sm];
    m /= (double)A.ne;
    const double dep = A.dat[n] - m;
    A.tmp_dep[n] = dep;
    A.tmp_mean[n] = m;
    bool ok = (A.qc == nullptr) || (A.qc[n] == 1);
    if (ok && fabs(dep) > A.gross_error * A.err[n]) ok = false;
    if (ok) {
        const double phi = A.olat[n], lam = A.olon[n];
        // lat band j: lat[j] <= phi < lat[j+1], j in 0..nlat-2
        int lo = -1;
        if (!(phi < A.lat[0]) && phi < A.lat[A.nlat - 1]) {
            int a = 0, b = A.nlat - 1;  // lat[a] <= phi < lat[b]
            while (b - a > 1) {
                const int c = (a + b) >> 1;
                if (A.lat[c] <= phi)
                    a = c;
                else
                    b = c;
            }
            lo = a;
        }
        int li = -1;
        if (!(lam < A.lon[0]) && lam < 360.0) {
            if (!(lam < A.lon[A.nlon - 1])) {
                li = A.nlon - 1;
            } else {
                int a = 0, b = A.nlon - 1;
                while (b - a > 1) {
                    const int c = (a + b) >> 1;
                    if (A.lon[c] <= lam)
                        a = c;
                    else
                        b = c;
                }
                li = a;
            }
        }
        if (lo >= 0 && li >= 0) {
            const unsigned cell = (unsigned)(lo * A.nlon + li);
            key = ((unsigned long long)cell << 32) | (unsigned)n;
            atomicAdd(&A.cell_cnt[cell], 1);
        }
    }
    A.keys[n] = key;
}

// bitonic sort steps on unique 64-bit keys (array length a power of two)
__global__ void bitonic_step(unsigned long long* keys, long n, int j, int kk)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long ixj = i ^ j;
    if (ixj > i) {
        const unsigned long long a = keys[i], b = keys[ixj];
        const bool up = ((i & kk) == 0);
        if ((a > b) == up) {
            keys[i] = b;
            keys[ixj] = a;
        }
    }
}

__global__ __launch_bounds__(1024) void bitonic_local(unsigned long long* keys, long n, int kk_max)
{
    // sorts runs of 2048 with all steps j < 1024 done in LDS
    __shared__ unsigned long long s[2048];
    const long base = (long)blockIdx.x * 2048;
    for (int t = threadIdx.x; t < 2048; t += 1024) s[t] = base + t < n ? keys[base + t] : ~0ull;
    __syncthreads();
    for (int kk = 2; kk <= kk_max; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < 2048; t += 1024) {
                const int ixj = t ^ j;
                if (ixj > t) {
                    const long gi = base + t;
                    const bool up = ((gi & kk) == 0);
                    const unsigned long long a = s[t], b = s[ixj];
                    if ((a > b) == up) {
                        s[t] = b;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < 2048; t += 1024)
        if (base + t < n) keys[base + t] = s[t];
}

__global__ __launch_bounds__(1024) void bitonic_merge_local(unsigned long long* keys, long n, int kk)
{
    // finishes stage kk for the steps j < 2048 in LDS
    __shared__ unsigned long long s[2048];
    const long base = (long)blockIdx.x * 2048;
    for (int t = threadIdx.x; t < 2048; t += 1024) s[t] = keys[base + t];
    __syncthreads();
    for (int j = 1024; j > 0; j >>= 1) {
        for (int t = threadIdx.x; t < 2048; t += 1024) {
            const int ixj = t ^ j;
            if (ixj > t) {
                const long gi = base + t;
                const bool up = ((gi & kk) == 0);
                const unsigned long long a = s[t], b = s[ixj];
                if ((a > b) == up) {
                    s[t] = b;
                    s[ixj] = a;
                }
            }
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < 2048; t += 1024) keys[base + t] = s[t];
}

// scatter the sorted obs into the binned arrays and precompute per-obs terms
__global__ void obs_scatter(ObsPrepArgs A, int nbinned)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nbinned) return;
    const int n = (int)(A.keys[m] & 0xffffffffull);
    A.perm[m] = n;
    const double mean = A.tmp_mean[n];
    for (int i = 0; i < A.ne; ++i) {
        const double v = A.hxf[(size_t)n * A.h_so + (size_t)i * A.h_sm] - mean;
        A.o_hdxf[(size_t)m * A.ne + i] = v;
        if (A.o_hdxf32) A.o_hdxf32[(size_t)m * A.ne + i] = (float)v;
    }
    A.o_dep[m] = A.tmp_dep[n];
    A.o_rdiag[m] = A.err[n] * A.err[n];
    A.o_err[m] = A.err[n];
    const int e = (int)llround(A.elm[n]);
    int typ = -1;
    switch (e) {
    case LEKF_ID_U: typ = 0; break;
    case LEKF_ID_V: typ = 1; break;
    case LEKF_ID_T: typ = 2; break;
    case LEKF_ID_Q: typ = 3; break;
    case LEKF_ID_RH: typ = 4; break;
    case LEKF_ID_PS: typ = 5; break;
    case LEKF_ID_RAIN: typ = 6; break;
    default: typ = -1;
    }
    A.o_typ[m] = typ;
    const double r180 = 1.0 / 180.0;
    const double latr = rn_mul(rn_mul(A.olat[n], LEKF_PI_REF), r180);
    A.o_sinlat[m] = sin(latr);
    A.o_coslat[m] = cos(latr);
    A.o_lonr[m] = rn_mul(rn_mul(A.olon[n], LEKF_PI_REF), r180);
    A.o_lnlev[m] = (typ == 5) ? 0.0 : log(A.olev[n]);
    A.o_lndat[m] = (typ == 5) ? log(A.dat[n]) : 0.0;
}

// nobsgrd(i, j) = cumulative count through cell (i, j) in (j, i) order
__global__ __launch_bounds__(1024) void cell_scan(const int* cell_cnt, int* nobsgrd, int ncell)
{
    __shared__ int s[1024];
    int carry = 0;
    for (int base = 0; base < ncell; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < ncell ? cell_cnt[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            const int t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < ncell) nobsgrd[i] = carry + s[threadIdx.x];
        const int tot = s[1023];
        __syncthreads();
        carry += tot;
    }
}

// ln(ps_mean * sig) per (level, column) from the resident ensemble
// (speedy/letkf/letkf_tools.f90:144-146 with calc_pfull)
__global__ void logpfm_kernel(const double* x2b, int nij, int nv2d, int ips, int ne,
                              const double* sig, int nlev, double* logpfm)
{
    const int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= nij) return;
    const double* row = x2b + ((size_t)ij * nv2d + ips) * ne;
    double s = row[0];
    for (int m = 1; m < ne; ++m) s += row[m];
    s /= (double)ne;
    for (int l = 0; l < nlev; ++l) logpfm[(size_t)l * nij + ij] = log(s * sig[l]);
}

}  // namespace

// Level groups of a fill launch over ncols columns: enough workgroups to give
// every CU several waves (a column chunk of a capped global grid holds only a
// few hundred columns), each group scanning the column's candidates for its
// own run of levels.  A level's list is filled in scan order whatever the
// grouping, so the lists do not depend on it.  4096 workgroups: on the
// GFS-sized grid (chunks of ~270 columns, 64 levels) the selection took 615 ms
// per cycle with 2048, 583 ms with 4096 and 615 ms with 8192.
static unsigned fill_level_groups(int ncols, int nlev)
{
    constexpr int kFillBlocks = 4096;
    const int g = (kFillBlocks + ncols - 1) / ncols;
    return (unsigned)(g < 1 ? 1 : (g > nlev ? nlev : g));
}

hipError_t lekf_cull_global_launch(const GlobalCullArgs& A, hipStream_t st)
{
    const int nb = (A.nij + 255) / 256;
    KTime kt("cull", st);
    hipLaunchKernelGGL(cull_global_count, dim3(nb), dim3(256), 0, st, A);
    hipLaunchKernelGGL(scan_ex, dim3(1), dim3(1024), 0, st, A.cand_cnt, A.cand_off, A.nij);
    hipLaunchKernelGGL(cull_global_fill, dim3(A.nij, fill_level_groups(A.nij, A.nlev)), dim3(64),
                       0, st, A);
    return hipGetLastError();
}

hipError_t lekf_cull_global_count_launch(const GlobalCullArgs& A, hipStream_t st)
{
    const int nb = (A.nij + 255) / 256;
    hipLaunchKernelGGL(cull_global_count, dim3(nb), dim3(256), 0, st, A);
    hipLaunchKernelGGL(scan_ex, dim3(1), dim3(1024), 0, st, A.cand_cnt, A.cand_off, A.nij);
    return hipGetLastError();
}

hipError_t lekf_cull_global_fill_launch(const GlobalCullArgs& A, int ncols, hipStream_t st)
{
    if (ncols <= 0) return hipSuccess;
    KTime kt("cull", st);
    hipLaunchKernelGGL(cull_global_fill, dim3(ncols, fill_level_groups(ncols, A.nlev)), dim3(64), 0,
                       st, A);
    return hipGetLastError();
}

hipError_t lekf_cull_l96_launch(const L96CullArgs& A, hipStream_t st)
{
    hipLaunchKernelGGL(cull_l96, dim3(A.nx), dim3(64), 0, st, A);
    return hipGetLastError();
}

hipError_t lekf_obs_prep_launch(const ObsPrepArgs& A, hipStream_t st)
{
    const int nb = (A.nobs + 255) / 256;
    hipLaunchKernelGGL(obs_prep, dim3(nb > 0 ? nb : 1), dim3(256), 0, st, A);
    return hipGetLastError();
}

hipError_t lekf_sort_keys_launch(unsigned long long* keys, long n2, hipStream_t st)
{
    // n2: power of two >= 2048
    const long nblk = n2 / 2048;
    hipLaunchKernelGGL(bitonic_local, dim3((unsigned)nblk), dim3(1024), 0, st, keys, n2, 2048);
    for (long kk = 4096; kk <= n2; kk <<= 1) {
        for (long j = kk >> 1; j >= 2048; j >>= 1)
            hipLaunchKernelGGL(bitonic_step, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st,
                               keys, n2, (int)j, (int)kk);
        hipLaunchKernelGGL(bitonic_merge_local, dim3((unsigned)nblk), dim3(1024), 0, st, keys, n2,
                           (int)kk);
    }
    return hipGetLastError();
}

hipError_t lekf_obs_scatter_launch(const ObsPrepArgs& A, int nbinned, int ncell, hipStream_t st)
{
    const int nb = (nbinned + 255) / 256;
    if (nbinned > 0)
        hipLaunchKernelGGL(obs_scatter, dim3(nb), dim3(256), 0, st, A, nbinned);
    hipLaunchKernelGGL(cell_scan, dim3(1), dim3(1024), 0, st, A.cell_cnt, A.nobsgrd, ncell);
    return hipGetLastError();
}

hipError_t lekf_logpfm_launch(const double* x2b, int nij, int nv2d, int ips, int ne,
                              const double* sig, int nlev, double* logpfm, hipStream_t st)
{
    hipLaunchKernelGGL(logpfm_kernel, dim3((nij + 255) / 256), dim3(256), 0, st, x2b, nij, nv2d,
                       ips, ne, sig, nlev, logpfm);
    return hipGetLastError();
}
