// lekf_radar.hip -- SCALE's radar obs operator on the resident ensemble (gfx950).
//
// For every (obs, member): the height and range checks of SCALE's obsope
// (scale/obs/obsope_tools.f90:353-361: RADAR_ZMAX, then phys2ijkz), the
// trilinear interpolation of u, v, w, T, p, q, qc, qr, qi, qs, qg at the obs
// (itpl_3d), and Trans_XtoY_radar / calc_ref_vr
// (scale/common/common_obs_scale.f90:338-476, 609-960): reflectivity in dBZ or
// radial velocity.  obsope turns iqc_ref_low back into good
// (obsope_tools.f90:380), and so does this kernel.
//
// Grid: nx x ny columns of nlev levels without halo; ri, rj, rk are 1-based
// grid coordinates.  The reference's halo lets it take ri = 1 (or rj = 1)
// and a height equal to the top level's; here the interpolation corners must
// lie in the grid: 1 < ri <= nx, 1 < rj <= ny (else out_h), and a height equal
// to the top level's takes the top layer.  Heights must ascend with the level
// (the reference assumes it); a column where they do not yields out_vlo.
//
// One wave per obs, lane = member: the position work is the same in every
// lane; a member's corner reads are coalesced across the wave.  Every
// arithmetic operation is rounded as written (the reference has no fused
// multiply-add), so the grid indices and rk are bit-identical to the
// reference's; the dBZ and velocity values go through the device pow, log10,
// atan2, acos, sin and cos.
#include "lekf_dev.hpp"
#include "lekf_radar.hpp"

#pragma clang fp contract(off)

namespace {

constexpr double kPi = LEKF_PI_REF;  // common/common.f90:28
constexpr double kRe = LEKF_RE_REF;
constexpr double kGg = 9.81;         // common/common.f90:29
constexpr double kRd = 287.05;       // common/common.f90:30
constexpr double kUndef = -9.99e33;  // common/common.f90:38

// default-real (single precision) literals of calc_ref_vr, promoted to double
__device__ __forceinline__ double f4(float x) { return (double)x; }

// itpl_3d with the level first, as Trans_XtoY_radar calls it: layer kl-1, kl
// (1-based), columns c00.. of the lower-left corner
__device__ __forceinline__ double itpl(const RadarArgs& A, int v, int m, int kl, long c00,
                                       double al, double ax, double ay)
{
    const long nij = (long)A.nx * A.ny;
    const long c10 = c00 + 1, c01 = c00 + A.nx, c11 = c01 + 1;
    auto x = [&](int lev1, long c) {  // lev1: 1-based level
        return A.x3[(((long)(lev1 - 1) * nij + c) * A.nv3d + v) * A.ne + m];
    };
    const double bl = 1.0 - al, bx = 1.0 - ax, by = 1.0 - ay;
    double s = x(kl - 1, c00) * bl * bx * by;
    s = s + x(kl, c00) * al * bx * by;
    s = s + x(kl - 1, c10) * bl * ax * by;
    s = s + x(kl, c10) * al * ax * by;
    s = s + x(kl - 1, c01) * bl * bx * ay;
    s = s + x(kl, c01) * al * bx * ay;
    s = s + x(kl - 1, c11) * bl * ax * ay;
    s = s + x(kl, c11) * al * ax * ay;
    return s;
}

// calc_ref_vr (common_obs_scale.f90:609-960)
__device__ void calc_ref_vr(const RadarArgs& A, double qr, double qs, double qg, double u,
                            double v, double w, double t, double p, double az, double elev,
                            double& ref_out, double& vr_out)
{
    const double pi = kPi, deg2rad = kPi / 180.0;
    double zr = 0.0, zs = 0.0, zg = 0.0, zms = 0.0, zmg = 0.0, ref = 0.0, wt = 0.0;
    double ro = p / (kRd * t);
    const double* gam = A.gam;
    if (A.method == 1) {
        const double nor = 8.0e6, ror = 1000.0, pip = pow(pi, 1.75), cf = 10.0e18 * 72, p0 = 1.0e5;
        const double qt = qr + qs + qg;
        if (qt > 0.0) {
            ref = cf * pow(ro * qt, 1.75);
            ref = ref / (pip * pow(nor, 0.75) * pow(ror, 1.75));
            const double a = pow(p0 / p, f4(0.4f));
            wt = 5.40 * a * pow(qt, 0.125);
        }
    } else if (A.method == 2) {
        double nor = 8.0e6, nos = 3.0e6, nog = 4.0e4, ror = 1000.0, ros = 100.0, rog = 913.0,
               roi = 917.0, roo = 1.0;
        const double ki2 = 0.176, kr2 = 0.930, pip = pow(pi, 1.75), cf = 1.0e18 * 720;
        if (qr > 0.0) {
            zr = cf * pow(ro * qr, 1.75);
            zr = zr / (pip * pow(nor, 0.75) * pow(ror, 1.75));
        }
        if (qs > 0.0) {
            if (t <= f4(273.16f)) {
                zs = cf * ki2 * pow(ros, 0.25) * pow(ro * qs, 1.75);
                zs = zs / (pip * kr2 * pow(nos, 0.75) * (roi * roi));
            } else {
                zs = cf * pow(ro * qs, 1.75);
                zs = zs / (pip * pow(nos, 0.75) * pow(roi, 1.75));
            }
        }
        if (qg > 0.0) {
            zg = pow(cf / (pip * pow(nog, 0.75) * pow(rog, 1.75)), f4(0.95f));
            zg = zg * pow(ro * qg, f4(1.6625f));
        }
        ref = zr + zs + zg;
        if (ref > 0.0) {
            const double m3 = f4(1e-3f);
            nor = nor * m3;
            nos = nos * m3;
            nog = nog * m3;
            ror = ror * m3;
            ros = ros * m3;
            rog = rog * m3;
            roo = roo * m3;
            ro = ro * m3;
            const double a = 2115.0, b = 0.8, c = 152.93, d = 0.25, Cd = 0.6;
            const double lr = pow(pi * ror * nor / (ro * qr), 0.25);
            const double ls = pow(pi * ros * nos / (ro * qs), 0.25);
            const double lg = pow(pi * rog * nog / (ro * qg), 0.25);
            const double rofactor = pow(roo / ro, 0.25);
            double wr = a * gam[0] / (6.0 * pow(lr, b));
            wr = 1.0e-2 * wr * rofactor;
            double ws = c * gam[1] / (6.0 * pow(ls, d));
            ws = 1.0e-2 * ws * rofactor;
            double wg = gam[2] * pow((4.0 * kGg * 100.0 * rog) / (3.0 * Cd * ro), 0.5);
            wg = 1.0e-2 * wg / (6.0 * pow(lg, 0.5));
            wt = (wr * zr + ws * zs + wg * zg) / (zr + zs + zg);
        }
    } else {  // method 3 (the launcher admits 1, 2, 3 only)
        const double maxf = 0.5;
        double Fg = 0.0, Fs = 0.0, fwg = 0.0, fws = 0.0;
        if (qr > 0.0 && qg > 0.0) {
            Fg = maxf * pow(fmin(qr / qg, qg / qr), 1.0 / 3.0);
            fwg = qr / (qr + qg);
        }
        if (qr > 0.0 && qs > 0.0) {
            Fs = maxf * pow(fmin(qr / qs, qs / qr), 1.0 / 3.0);
            fws = qr / (qr + qs);
        }
        const double qrp = (1.0 - Fs - Fg) * qr, qsp = (1.0 - Fs) * qs, qgp = (1.0 - Fg) * qg;
        const double qms = Fs * (qr + qs), qmg = Fg * (qr + qg);
        if (qrp > 0.0) zr = 2.53e4 * pow(ro * qrp * 1.0e3, f4(1.84f));
        if (qsp > 0.0) zs = 3.48e3 * pow(ro * qsp * 1.0e3, f4(1.66f));
        if (qgp > 0.0) zg = 8.18e4 * pow(ro * qgp * 1.0e3, 1.50);
        if (qms > 0.0) {
            zms = (f4(0.00491f) + 5.75 * fws - f4(5.588f) * (fws * fws)) * 1.0e5;
            zms = zms * pow(ro * qms * 1.0e3,
                            f4(1.67f) - f4(0.202f) * fws + f4(0.398f) * (fws * fws));
        }
        if (qmg > 0.0) {
            zmg = (f4(0.809f) + f4(10.13f) * fwg - f4(5.98f) * (fwg * fwg)) * 1.0e5;
            zmg = zmg * pow(ro * qmg * 1.0e3,
                            f4(1.48f) + f4(0.0448f) * fwg - f4(0.0313f) * (fwg * fwg));
        }
        ref = zr + zg + zs + zms + zmg;
        if (ref > 0.0) {
            const double nor = 8.0e-2, nos = 3.0e-2, nog = 4.0e-4, ror = 1.0, ros = 0.1,
                         rog = 0.917, roo = 0.001;
            ro = 1.0e-3 * ro;
            const double a = 2115.0, b = 0.8, c = 152.93, d = 0.25, Cd = 0.6;
            const double rofactor = pow(roo / ro, 0.5);
            double wr = 0.0, ws = 0.0, wg = 0.0;
            if (qr > 0.0) {
                const double lr = pow(pi * ror * nor / (ro * qr), 0.25);
                wr = a * gam[0] / (6.0 * pow(lr, b));
                wr = 1.0e-2 * wr * rofactor;
            }
            if (qs > 0.0) {
                const double ls = pow(pi * ros * nos / (ro * qs), 0.25);
                ws = c * gam[1] / (6.0 * pow(ls, d));
                ws = 1.0e-2 * ws * rofactor;
            }
            if (qg > 0.0) {
                const double lg = pow(pi * rog * nog / (ro * qg), 0.25);
                wg = gam[2] * pow((4.0 * kGg * 100.0 * rog) / (3.0 * Cd * ro), 0.5);
                wg = 1.0e-2 * wg / (6.0 * pow(lg, 0.5));
            }
            wt = (wr * zr + ws * zs + ws * zms + wg * zg + wg * zmg) / (zr + zs + zg + zms + zmg);
        }
    }
    double vr = u * cos(elev * deg2rad) * sin(az * deg2rad);
    vr = vr + v * cos(elev * deg2rad) * cos(az * deg2rad);
    if (A.use_wt)
        vr = vr + (w - wt) * sin(elev * deg2rad);
    else
        vr = vr + w * sin(elev * deg2rad);
    ref_out = ref;
    vr_out = vr;
}

__global__ __launch_bounds__(64) void radar_obsope(RadarArgs A)
{
    const int n = blockIdx.x;
    const int lane = threadIdx.x;
    const long nij = (long)A.nx * A.ny;
    const int elm = (int)A.elm[n];
    const double lon = A.lon[n], lat = A.lat[n], lev = A.z[n], ri = A.ri[n], rj = A.rj[n];
    int qc = LEKF_IQC_GOOD;
    double rk = kUndef;
    int i = 0, j = 0, kl = 0;
    double ai = 0.0, aj = 0.0;
    if (elm != LEKF_ID_RADAR_REF && elm != LEKF_ID_RADAR_VR) {
        qc = LEKF_IQC_OTYPE;
    } else if (lev > A.zmax) {
        qc = LEKF_IQC_RADAR_VHI;
    } else if (!(ri > 1.0 && ri <= (double)A.nx && rj > 1.0 && rj <= (double)A.ny)) {
        qc = LEKF_IQC_OUT_H;
    } else {
        // phys2ijkz (common_obs_scale.f90): the lowest level with a valid
        // height in all four corner columns, the column heights at (ri, rj),
        // and the bracketing layer
        i = (int)ceil(ri);
        j = (int)ceil(rj);
        ai = ri - (double)(i - 1);
        aj = rj - (double)(j - 1);
        const long c00 = (i - 2) + (long)A.nx * (j - 2);
        int ks = 1;
        for (int jj = 0; jj < 2; ++jj)
            for (int ii = 0; ii < 2; ++ii) {
                const long c = c00 + ii + (long)A.nx * jj;
                int k = 1;
                for (; k <= A.nlev; ++k) {
                    const double z = A.hgt[c + nij * (k - 1)];
                    if (z > -300.0 && z < 10000.0) break;
                }
                if (k > ks) ks = k;
            }
        auto zcol = [&](int k1) {
            const double* z = A.hgt + nij * (k1 - 1);
            const double bi = 1.0 - ai, bj = 1.0 - aj;
            double s = z[c00] * bi * bj;
            s = s + z[c00 + 1] * ai * bj;
            s = s + z[c00 + A.nx] * bi * aj;
            s = s + z[c00 + A.nx + 1] * ai * aj;
            return s;
        };
        if (ks > A.nlev) {
            qc = LEKF_IQC_OUT_VLO;
        } else if (lev > zcol(A.nlev)) {
            qc = LEKF_IQC_OUT_VHI;
        } else if (lev < zcol(ks)) {
            qc = LEKF_IQC_OUT_VLO;
        } else {
            int k = ks + 1;
            double zk = 0.0;
            for (; k <= A.nlev; ++k) {
                zk = zcol(k);
                if (zk > lev) break;
            }
            if (k > A.nlev) {
                k = A.nlev;
                zk = zcol(k);
            }
            const double zk1 = zcol(k - 1);
            const double ak = (lev - zk1) / (zk - zk1);
            rk = (double)(k - 1) + ak;
            if (!(rk >= 1.0 && rk <= (double)A.nlev)) {
                // heights not ascending at this column: no layer to interpolate in
                qc = LEKF_IQC_OUT_VLO;
                rk = kUndef;
            } else {
                kl = (int)ceil(rk);
                if (kl < 2) kl = 2;
            }
        }
    }
    double az = 0.0, elev = 0.0;
    if (qc == LEKF_IQC_GOOD) {
        const double deg2rad = kPi / 180.0, rad2deg = 180.0 / kPi;
        const double dlon = lon - A.rlon[n], dlat = lat - A.rlat[n];
        if (dlon == 0.0 && dlat == 0.0) {
            qc = LEKF_IQC_OUT_H;
        } else {
            az = rad2deg * atan2(dlon * cos(A.rlat[n] * deg2rad), dlat);
            if (az < 0.0) az = 360.0 + az;
            // com_distll_1 (common/common.f90:401-424)
            const double r180 = 1.0 / 180.0;
            const double lon1 = lon * kPi * r180, lon2 = A.rlon[n] * kPi * r180;
            const double lat1 = lat * kPi * r180, lat2 = A.rlat[n] * kPi * r180;
            double cosd = sin(lat1) * sin(lat2) + cos(lat1) * cos(lat2) * cos(lon2 - lon1);
            cosd = fmin(1.0, cosd);
            cosd = fmax(-1.0, cosd);
            const double dist = acos(cosd) * kRe;
            elev = rad2deg * atan2(lev - A.rz[n], dist);
        }
    }
    if (lane == 0) A.rk[n] = rk;
    for (int m = lane; m < A.ne; m += LEKF_WAVE) {
        double y = kUndef;
        if (qc == LEKF_IQC_GOOD) {
            const long c00 = (i - 2) + (long)A.nx * (j - 2);
            const double al = rk - (double)(kl - 1);
            double f[11];
#pragma unroll
            for (int q = 0; q < 11; ++q) f[q] = itpl(A, A.var[q], m, kl, c00, al, ai, aj);
            double ref = 0.0, vr = 0.0;
            calc_ref_vr(A, f[7], f[9], f[10], f[0], f[1], f[2], f[3], f[4], az, elev, ref, vr);
            if (elm == LEKF_ID_RADAR_REF)
                y = ref < pow(10.0, A.min_ref_dbz / 10.0) ? A.min_ref_dbz + A.low_ref_shift
                                                          : 10.0 * log10(ref);
            else
                y = vr;
        }
        A.hxf[(size_t)n * A.ne + m] = y;
        A.qc[(size_t)n * A.ne + m] = qc;
    }
}

// SCALE's departure and QC step (scale/letkf/letkf_obs.f90:320-343 obs
// preprocessing, :357 the merged operator qc, :476-508 radar checks, :560-576
// mean / perturbations / departure, :590-645 gross error).  One thread per
// obs; member m of obs n at [n * so + m * sm]: column-major (so = 1, sm =
// nobs) a wave reads 64 consecutive obs of one member per load, the radar
// operator's rows (so = ne, sm = 1) are read in place.  The mean is summed in
// member order as the reference does.
__global__ __launch_bounds__(256) void scale_obs_qc(ScaleQcArgs A)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= A.nobs) return;
    const long so = A.so, sm = A.sm;
    const int elm = (int)A.elm[n];
    double y = A.dat_in[n], e = A.err_in[n];
    if (elm == LEKF_ID_RADAR_REF) {
        if (y >= 0.0 && y < 1.0e10)
            y = y < A.min_ref ? A.min_ref_dbz + A.low_ref_shift : 10.0 * log10(y);
        else
            y = LEKF_UNDEF;
        if (A.use_obserr_ref) e = A.obserr_ref;
    }
    if (elm == LEKF_ID_RADAR_VR && A.use_obserr_vr) e = A.obserr_vr;
    A.dat[n] = y;
    A.err[n] = e;
    A.dep[n] = 0.0;
    int qc = LEKF_IQC_GOOD;
    if (A.qc_mem)
        for (int m = 0; m < A.ne; ++m) qc = max(qc, A.qc_mem[n * so + m * sm]);
    if (qc == LEKF_IQC_GOOD && elm == LEKF_ID_RADAR_REF) {
        if (!A.use_ref) {
            qc = LEKF_IQC_OTYPE;
        } else if (y == LEKF_UNDEF) {
            qc = LEKF_IQC_OBS_BAD;
        } else {
            const double t = A.thres + 1.0e-6;
            int mem_ref = 0;
            for (int m = 0; m < A.ne; ++m) mem_ref += A.hx[n * so + m * sm] > t ? 1 : 0;
            if (mem_ref < (y > t ? A.min_member_obsref : A.min_member)) qc = LEKF_IQC_REF_MEM;
        }
    }
    if (qc == LEKF_IQC_GOOD && elm == LEKF_ID_RADAR_VR && !A.use_vr) qc = LEKF_IQC_OTYPE;
    if (qc != LEKF_IQC_GOOD) {
        A.qc[n] = qc;
        return;
    }
    double mean = A.hx[n * so];
    for (int m = 1; m < A.ne; ++m) mean = mean + A.hx[n * so + m * sm];
    mean = mean / (double)A.ne;
    for (int m = 0; m < A.ne; ++m) A.hx[n * so + m * sm] = A.hx[n * so + m * sm] - mean;
    const double d = y - mean;
    A.dep[n] = d;
    double fac;
    switch (elm) {
    case LEKF_ID_SCALE_RAIN: fac = A.gross_rain; break;
    case LEKF_ID_RADAR_REF: fac = A.gross_ref; break;
    case LEKF_ID_RADAR_VR: fac = A.gross_vr; break;
    case LEKF_ID_RADAR_PRH: fac = A.gross_prh; break;
    case LEKF_ID_TCLON: fac = A.gross_tcx; break;
    case LEKF_ID_TCLAT: fac = A.gross_tcy; break;
    case LEKF_ID_TCMIP: fac = A.gross_tcp; break;
    default: fac = A.gross; break;
    }
    A.qc[n] = fabs(d) > fac * e ? LEKF_IQC_GROSS_ERR : LEKF_IQC_GOOD;
}

}  // namespace

hipError_t lekf_radar_launch(const RadarArgs& A, hipStream_t st)
{
    if (A.nobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(radar_obsope, dim3((unsigned)A.nobs), dim3(LEKF_WAVE), 0, st, A);
    return hipGetLastError();
}

hipError_t lekf_scale_qc_launch(const ScaleQcArgs& A, hipStream_t st)
{
    if (A.nobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(scale_obs_qc, dim3((unsigned)((A.nobs + 255) / 256)), dim3(256), 0, st, A);
    return hipGetLastError();
}
