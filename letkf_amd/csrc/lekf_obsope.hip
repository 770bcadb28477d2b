// lekf_obsope.hip -- the global obs operator H(x) on the resident ensemble (gfx950).
//
// For every obs and member, the reference's operator chain
//   phys2ijk  (speedy/common/common_obs_speedy.f90:137-211): ri from the
//             longitude, rj piecewise-linear in the latitude table with the
//             poles extrapolated, rk linear in ln p between the member's
//             ln p_full levels interpolated to the obs column;
//   range and type rules of obsope (speedy/letkf/obsope.f90:56-89);
//   Trans_XtoY / itpl_2d / itpl_3d (common_obs_speedy.f90:30-81,215-282):
//             CEILING corners, longitude wrap to column 1, bi-/trilinear
//             weights in the reference's evaluation order;
// with p_full = ps * sig (speedy/common/common_speedy.f90:291-309).  Every
// operation is rounded as written (no contraction), so ri, rj and the corner
// indices are bit-identical to the reference; rk and H(x) pass through the
// device log, which may differ from the host's in the last place.
// An obs is accepted only if every member accepts it (the MINVAL of the
// member flags, speedy/letkf/letkf_obs.f90:160-161).
//
// One wave per obs, lane m = member m: the member's ensemble rows are
// contiguous in k, so the corner reads of a wave are coalesced.
#include "lekf_dev.hpp"
#include "lekf_obsope.hpp"

// Every FP64 operation in this file is rounded as written: no FMA contraction
// (the reference's results are reproduced operation by operation).
#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ double mul(double a, double b) { return rn_mul(a, b); }
__device__ __forceinline__ double add(double a, double b) { return rn_add(a, b); }
__device__ __forceinline__ double sub(double a, double b) { return rn_sub(a, b); }

// bilinear weights sum in the reference's order: v00 (1-ai)(1-aj) + v10 ai (1-aj)
// + v01 (1-ai) aj + v11 ai aj
__device__ __forceinline__ double bilin(double v00, double v10, double v01, double v11, double ai,
                                        double aj)
{
    const double bi = sub(1.0, ai), bj = sub(1.0, aj);
    double s = mul(mul(v00, bi), bj);
    s = add(s, mul(mul(v10, ai), bj));
    s = add(s, mul(mul(v01, bi), aj));
    s = add(s, mul(mul(v11, ai), aj));
    return s;
}

__global__ __launch_bounds__(64) void obsope_global(ObsOpArgs A)
{
    const int n = blockIdx.x;
    const int lane = threadIdx.x;
    const int nlon = A.nlon, nlat = A.nlat, nlev = A.nlev, ne = A.ne;
    const long nij = (long)nlon * nlat;
    const int elm = (int)rint(A.elm[n]);
    const double rlon = A.olon[n], rlat = A.olat[n], rlev = A.olev[n], odat = A.odat[n];
    bool ok = true;
    // longitude -> ri
    double ri;
    if (rlon == 0.0 || rlon == 360.0)
        ri = (double)(nlon + 1);
    else
        ri = add(mul(rn_div(rlon, 360.0), (double)nlon), 1.0);
    if (ceil(ri) < 2 || nlon + 1 < ceil(ri)) ok = false;
    // latitude -> rj
    double rj = 0.0;
    if (ok) {
        int j = 1;
        for (; j <= nlat; ++j)
            if (rlat < A.lat[j - 1]) break;
        if (j == 1) {
            rj = rn_div(add(rlat, 90.0), add(A.lat[0], 90.0));
        } else if (j == nlat + 1) {
            const double aj = rn_div(sub(rlat, A.lat[nlat - 1]), sub(90.0, A.lat[nlat - 1]));
            rj = add((double)nlat, aj);
        } else {
            const double aj = rn_div(sub(rlat, A.lat[j - 2]), sub(A.lat[j - 1], A.lat[j - 2]));
            rj = add((double)(j - 1), aj);
        }
        if (ceil(rj) < 2 || nlat < ceil(rj)) ok = false;
    }
    const bool known = elm == LEKF_ID_U || elm == LEKF_ID_V || elm == LEKF_ID_T ||
                       elm == LEKF_ID_Q || elm == LEKF_ID_PS;
    if (!known) ok = false;
    if (elm == LEKF_ID_PS && odat < -100.0) ok = false;
    // corners (1-based i, j as in the reference; ie wraps to column 1)
    const int i = ok ? (int)ceil(ri) : 2, j = ok ? (int)ceil(rj) : 2;
    const double ai = ok ? sub(ri, (double)(i - 1)) : 0.0, aj = ok ? sub(rj, (double)(j - 1)) : 0.0;
    const int ie = i <= nlon ? i : 1;
    const long c00 = (i - 2) + (long)nlon * (j - 2), c10 = (ie - 1) + (long)nlon * (j - 2);
    const long c01 = (i - 2) + (long)nlon * (j - 1), c11 = (ie - 1) + (long)nlon * (j - 1);
    for (int m = lane; m < ne; m += 64) {
        bool okm = ok;
        double y = 0.0, rk = 0.0;
        if (okm) {
            const double* ps = A.x2 + (size_t)A.ips * ne + m;  // row (ij * nv2d + ips)
            const size_t s2 = (size_t)A.nv2d * ne;
            const double p00 = ps[c00 * s2], p10 = ps[c10 * s2], p01 = ps[c01 * s2],
                         p11 = ps[c11 * s2];
            if (elm != LEKF_ID_PS) {
                // ln p_full at the obs column, level by level, and the bracketing level
                auto plev = [&](int k) {  // 1-based level
                    const double sg = A.sig[k - 1];
                    return bilin(log(mul(p00, sg)), log(mul(p10, sg)), log(mul(p01, sg)),
                                 log(mul(p11, sg)), ai, aj);
                };
                const double lr = log(rlev);
                double pk1 = plev(1), pk = 0.0;
                int k = 2;
                for (; k <= nlev - 1; ++k) {
                    pk = plev(k);
                    if (pk < lr) break;
                    pk1 = pk;
                }
                if (k == nlev) pk = plev(nlev);
                const double ak = rn_div(sub(lr, pk1), sub(pk, pk1));
                rk = add((double)(k - 1), ak);
                if (ceil(rk) > nlev) okm = false;
                if (okm && ceil(rk) < 2) {
                    if (elm == LEKF_ID_U || elm == LEKF_ID_V)
                        rk = 1.00001;
                    else
                        okm = false;
                }
            }
            if (okm) {
                if (elm == LEKF_ID_PS) {
                    y = bilin(p00, p10, p01, p11, ai, aj);
                } else {
                    const int v = elm == LEKF_ID_U ? A.iu : elm == LEKF_ID_V ? A.iv
                                : elm == LEKF_ID_T ? A.it : A.iq;
                    const int k = (int)ceil(rk);
                    const double ak = sub(rk, (double)(k - 1));
                    const size_t s3 = (size_t)A.nv3d * ne;
                    const double* x = A.x3 + (size_t)v * ne + m;
                    const long lo = (long)(k - 2) * nij, hi = (long)(k - 1) * nij;
                    const double bk = sub(1.0, ak);
                    const double bi = sub(1.0, ai), bj = sub(1.0, aj);
                    // itpl_3d order: (k-1 plane: 00, 10, 01, 11), then the k plane
                    double s = mul(mul(mul(x[(lo + c00) * s3], bi), bj), bk);
                    s = add(s, mul(mul(mul(x[(lo + c10) * s3], ai), bj), bk));
                    s = add(s, mul(mul(mul(x[(lo + c01) * s3], bi), aj), bk));
                    s = add(s, mul(mul(mul(x[(lo + c11) * s3], ai), aj), bk));
                    s = add(s, mul(mul(mul(x[(hi + c00) * s3], bi), bj), ak));
                    s = add(s, mul(mul(mul(x[(hi + c10) * s3], ai), bj), ak));
                    s = add(s, mul(mul(mul(x[(hi + c01) * s3], bi), aj), ak));
                    s = add(s, mul(mul(mul(x[(hi + c11) * s3], ai), aj), ak));
                    y = s;
                }
            }
        }
        A.hxf[(size_t)n * ne + m] = okm ? y : 0.0;
        if (A.rk) A.rk[(size_t)n * ne + m] = okm ? rk : 0.0;
        if (!okm) A.qc[n] = 0;  // several lanes may store the same 0
    }
    if (lane == 0) {
        if (A.ri) A.ri[n] = ok ? ri : 0.0;
        if (A.rj) A.rj[n] = ok ? rj : 0.0;
    }
}

__global__ void fill_int(int* a, int v, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

}  // namespace

hipError_t lekf_obsope_global_launch(const ObsOpArgs& A, hipStream_t st)
{
    if (A.nobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_int, dim3((A.nobs + 255) / 256), dim3(256), 0, st, A.qc, 1, A.nobs);
    hipLaunchKernelGGL(obsope_global, dim3(A.nobs), dim3(64), 0, st, A);
    return hipGetLastError();
}
