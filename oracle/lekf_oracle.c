/*
 * lekf_oracle.c -- CPU restatement of the reference LETKF local analysis.
 *
 * TEST INFRASTRUCTURE (the oracle).  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py load this library, and only as the checker
 * or as the timed CPU baseline.  The product (letkf_amd/) never links, loads or
 * calls it.
 *
 * Each function names the reference file:line it restates.  Matrices use the
 * reference's column-major layout (element (i,j) of an ld x n matrix is at
 * a[i + j*ld], 0-based).
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   orc_letkf_core / orc_mtx_eigen / orc_tred2 / orc_tql2
 *       -> reference Fortran built from /root/reference (oracle/_ref), golden
 *          vectors under tests/golden/, closed forms (p=0, p=1 Sherman-Morrison),
 *          dense Kalman-filter identity, numpy.linalg.eigh.
 *   orc_obs_local_global / orc_obs_bins_global
 *       -> brute-force scan over all obs (set equality), wrap-around and pole cases.
 *   orc_obs_local_l96 -> closed form (p = 21 at every point for the L96 config).
 *   orc_obs_local_regional -> brute force + insertion-sort cap restated independently
 *          in the test.
 *   orc_phys2ij / orc_itpl_* / orc_trans_xtoy -> reproduction of fields linear in
 *          the grid indices, the longitude wrap to column 1, latitude-table
 *          points mapping to integer rj, the obsope vertical and type rules
 *          (tests/test_oracle_obsope.py).
 *   orc_l96_rk4 -> the reference model compiled from its sources
 *          (oracle/_ref, nx = 40) bit for bit, the fixed point x_i = F, and
 *          step-size convergence (tests/test_oracle_l96.py).
 *   orc_relax_point -> RTPP with alpha 0 / 1, RTPS restoring the background
 *          spread at alpha 1, beta = 0 giving the background, the spread/mean
 *          of the limited q row, and a numpy restatement
 *          (tests/test_oracle_relax.py).
 *   orc_radar_xtoy and its parts -> tests/test_oracle_radar.py.
 *   orc_obs_departure_scale -> known answers of each rule at and beside its
 *          threshold, the dBZ conversion at powers of ten, mean-free
 *          perturbations, the departure against a direct mean
 *          (tests/test_oracle_departure.py).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* common/common.f90:28,35 -- the truncated pi is part of the selection rule */
#define ORC_PI 3.1415926535
#define ORC_RE 6371.3e3

/* speedy/common/common_obs_speedy.f90:17-24 */
#define ID_U 2819
#define ID_V 2820
#define ID_T 3073
#define ID_Q 3330
#define ID_RH 3331
#define ID_PS 14593
#define ID_RAIN 9999

/* ------------------------------------------------------------------------- */
/* EISPACK restatement                                                        */
/* ------------------------------------------------------------------------- */

/* sqrt(a^2+b^2) by the Moler-Morrison iteration: common/netlib.f:504-523 */
double orc_pythag(double a, double b)
{
    double aa = fabs(a), ab = fabs(b);
    double p = aa > ab ? aa : ab;
    if (p == 0.0) return p;
    double q = (aa < ab ? aa : ab) / p;
    double r = q * q;
    for (;;) {
        double t = 4.0 + r;
        if (t == 4.0) break;
        double s = r / t;
        double u = 1.0 + 2.0 * s;
        p = u * p;
        r = (s / u) * (s / u) * r;
    }
    return p;
}

/*
 * Householder reduction of a symmetric matrix to tridiagonal form with the
 * transformation accumulated in z (Martin, Reinsch & Wilkinson 1968).
 * common/netlib.f:1023-1186 (tred2).  Only the lower triangle of a is read.
 * On return d = diagonal, e[1..n-1] = sub-diagonal, e[0] = 0.
 */
void orc_tred2(int n, const double *a, double *d, double *e, double *z)
{
#define Z(i, j) z[(size_t)(i) + (size_t)(j) * (size_t)n]
#define A(i, j) a[(size_t)(i) + (size_t)(j) * (size_t)n]
    for (int c = 0; c < n; ++c) {
        for (int r = c; r < n; ++r) Z(r, c) = A(r, c);
        d[c] = A(n - 1, c);
    }
    if (n == 1) goto finish;
    for (int i = n - 1; i >= 1; --i) {
        const int l = i - 1;
        double h = 0.0, scale = 0.0;
        if (l >= 1)
            for (int k = 0; k <= l; ++k) scale += fabs(d[k]);
        if (l < 1 || scale == 0.0) {
            e[i] = d[l];
            for (int j = 0; j <= l; ++j) {
                d[j] = Z(l, j);
                Z(i, j) = 0.0;
                Z(j, i) = 0.0;
            }
        } else {
            for (int k = 0; k <= l; ++k) {
                d[k] /= scale;
                h += d[k] * d[k];
            }
            double f = d[l];
            double g = -copysign(sqrt(h), f);
            e[i] = scale * g;
            h -= f * g;
            d[l] = f - g;
            for (int j = 0; j <= l; ++j) e[j] = 0.0;
            for (int j = 0; j <= l; ++j) {
                f = d[j];
                Z(j, i) = f;
                g = e[j] + Z(j, j) * f;
                for (int k = j + 1; k <= l; ++k) {
                    g += Z(k, j) * d[k];
                    e[k] += Z(k, j) * f;
                }
                e[j] = g;
            }
            f = 0.0;
            for (int j = 0; j <= l; ++j) {
                e[j] /= h;
                f += e[j] * d[j];
            }
            const double hh = f / (h + h);
            for (int j = 0; j <= l; ++j) e[j] -= hh * d[j];
            for (int j = 0; j <= l; ++j) {
                f = d[j];
                g = e[j];
                for (int k = j; k <= l; ++k) Z(k, j) = Z(k, j) - f * e[k] - g * d[k];
                d[j] = Z(l, j);
                Z(i, j) = 0.0;
            }
        }
        d[i] = h;
    }
    /* accumulate the transformations */
    for (int i = 1; i < n; ++i) {
        const int l = i - 1;
        Z(n - 1, l) = Z(l, l);
        Z(l, l) = 1.0;
        const double h = d[i];
        if (h != 0.0) {
            for (int k = 0; k <= l; ++k) d[k] = Z(k, i) / h;
            for (int j = 0; j <= l; ++j) {
                double g = 0.0;
                for (int k = 0; k <= l; ++k) g += Z(k, i) * Z(k, j);
                for (int k = 0; k <= l; ++k) Z(k, j) -= g * d[k];
            }
        }
        for (int k = 0; k <= l; ++k) Z(k, i) = 0.0;
    }
finish:
    for (int i = 0; i < n; ++i) {
        d[i] = Z(n - 1, i);
        Z(n - 1, i) = 0.0;
    }
    Z(n - 1, n - 1) = 1.0;
    e[0] = 0.0;
#undef Z
#undef A
}

/*
 * Implicit QL iteration on the tridiagonal (d, e) with the rotations applied
 * to z; eigenvalues returned ascending with matching columns of z.
 * common/netlib.f:718-887 (tql2).  Returns 0, or l+1 (1-based) when the l-th
 * eigenvalue needed more than 30 iterations.
 */
int orc_tql2(int n, double *d, double *e, double *z)
{
#define Z(i, j) z[(size_t)(i) + (size_t)(j) * (size_t)n]
    if (n == 1) return 0;
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    double f = 0.0, tst1 = 0.0;
    e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        double h = fabs(d[l]) + fabs(e[l]);
        if (tst1 < h) tst1 = h;
        int m = l;
        for (; m < n; ++m)
            if (tst1 + fabs(e[m]) == tst1) break;
        if (m != l) {
            double tst2;
            do {
                if (iter == 30) return l + 1;
                ++iter;
                const int l1 = l + 1, l2 = l + 2;
                double g = d[l];
                double p = (d[l1] - g) / (2.0 * e[l]);
                double r = orc_pythag(p, 1.0);
                d[l] = e[l] / (p + copysign(r, p));
                d[l1] = e[l] * (p + copysign(r, p));
                const double dl1 = d[l1];
                h = g - d[l];
                for (int i = l2; i < n; ++i) d[i] -= h;
                f += h;
                p = d[m];
                double c = 1.0, c2 = c, c3 = c;
                const double el1 = e[l1];
                double s = 0.0, s2 = 0.0;
                for (int i = m - 1; i >= l; --i) {
                    c3 = c2;
                    c2 = c;
                    s2 = s;
                    g = c * e[i];
                    h = c * p;
                    r = orc_pythag(p, e[i]);
                    e[i + 1] = s * r;
                    s = e[i] / r;
                    c = p / r;
                    p = c * d[i] - s * g;
                    d[i + 1] = h + s * (c * g + s * d[i]);
                    for (int k = 0; k < n; ++k) {
                        const double zh = Z(k, i + 1);
                        Z(k, i + 1) = s * Z(k, i) + c * zh;
                        Z(k, i) = c * Z(k, i) - s * zh;
                    }
                }
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p;
                d[l] = c * p;
                tst2 = tst1 + fabs(e[l]);
            } while (tst2 > tst1);
        }
        d[l] += f;
    }
    /* selection sort to ascending order, swapping vectors */
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        double p = d[i];
        for (int j = i + 1; j < n; ++j)
            if (d[j] < p) {
                k = j;
                p = d[j];
            }
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            for (int j = 0; j < n; ++j) {
                const double t = Z(j, i);
                Z(j, i) = Z(j, k);
                Z(j, k) = t;
            }
        }
    }
    return 0;
#undef Z
}

/*
 * mtx_eigen(imode=1): common/common_mtx.f90:41-99.  Eigenvalues descending,
 * eigenpairs with lambda < lambda_max*sqrt(eps) zeroed (rank truncation).
 * Returns 0, or 2 where the reference executes STOP 2.
 * The compaction block at common_mtx.f90:80-91 only runs when the smallest
 * eigenvalue survived truncation, which is impossible once any was truncated;
 * it is therefore a no-op and is not restated.
 */
int orc_mtx_eigen(int n, const double *a, double *eival, double *eivec, int *nrank_eff)
{
    double *z = (double *)calloc((size_t)n * n, sizeof(double));
    double *d = (double *)malloc(sizeof(double) * n);
    double *e = (double *)malloc(sizeof(double) * n);
    orc_tred2(n, a, d, e, z);
    int ierr = orc_tql2(n, d, e, z);
    int rc = 0;
    if (ierr != 0) {
        rc = 2;
        goto out;
    }
    int nrank = n;
    if (d[n - 1] > 0) {
        const double thr = fabs(d[n - 1]) * sqrt(DBL_EPSILON);
        for (int i = 0; i < n; ++i)
            if (d[i] < thr) {
                --nrank;
                d[i] = 0.0;
                for (int r = 0; r < n; ++r) z[r + (size_t)i * n] = 0.0;
            }
    } else {
        rc = 2;
        goto out;
    }
    for (int i = 0; i < n; ++i) {
        eival[i] = d[n - 1 - i];
        memcpy(eivec + (size_t)i * n, z + (size_t)(n - 1 - i) * n, sizeof(double) * n);
    }
    if (nrank_eff) *nrank_eff = nrank;
out:
    free(z);
    free(d);
    free(e);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* letkf_core: common/common_letkf.f90:50-243                                 */
/* ------------------------------------------------------------------------- */

/*
 * hdxb is ldh x ne column-major (first nobsl rows used).  transm / pao / minfl
 * may be NULL (the Fortran OPTIONAL arguments).  Returns 0, 2 (eigensolver
 * failure, the reference's STOP 2) or 99 (bad sizes).
 * The matrix products follow the reference DGEMM association and loop order
 * (common/netlibblas.f:243-555): C = Yr^T Y, Pa = (V/L) V^T, work2 = Pa Yr^T,
 * w = work2 d, T = (V sqrt((k-1)/L)) V^T.
 */
int orc_letkf_core(int ne, int ldh, int nobsl, const double *hdxb, const double *rdiag,
                   const double *rloc, const double *dep, double *parm_infl, double *trans,
                   double *transm, double *pao, int rdiag_wloc, const double *minfl)
{
    if (ne < 1 || nobsl < 0 || ldh < nobsl) return 99;
    const size_t k = (size_t)ne;
    if (nobsl == 0) { /* common_letkf.f90:82-96 */
        memset(trans, 0, sizeof(double) * k * k);
        for (size_t i = 0; i < k; ++i) trans[i + i * k] = sqrt(*parm_infl);
        if (transm) memset(transm, 0, sizeof(double) * k);
        if (pao) {
            memset(pao, 0, sizeof(double) * k * k);
            for (size_t i = 0; i < k; ++i) pao[i + i * k] = *parm_infl / (double)(ne - 1);
        }
        return 0;
    }
    const size_t p = (size_t)nobsl;
    double *yr = (double *)malloc(sizeof(double) * p * k);
    double *w1 = (double *)malloc(sizeof(double) * k * k);
    double *ev = (double *)malloc(sizeof(double) * k * k);
    double *el = (double *)malloc(sizeof(double) * k);
    double *pa = (double *)malloc(sizeof(double) * k * k);
    double *w2 = (double *)malloc(sizeof(double) * k * p);
    double *w3 = (double *)malloc(sizeof(double) * k);
    int rc = 0;
    /* :101-113 */
    for (size_t j = 0; j < k; ++j)
        for (size_t i = 0; i < p; ++i) {
            const double v = hdxb[i + j * (size_t)ldh] / rdiag[i];
            yr[i + j * p] = rdiag_wloc ? v : v * rloc[i];
        }
    /* :117-118 work1 = yr^T hdxb */
    for (size_t j = 0; j < k; ++j)
        for (size_t i = 0; i < k; ++i) {
            double t = 0.0;
            for (size_t l = 0; l < p; ++l) t += yr[l + i * p] * hdxb[l + j * (size_t)ldh];
            w1[i + j * k] = t;
        }
    /* :130-138 */
    if (minfl && *minfl > 0.0 && *parm_infl < *minfl) *parm_infl = *minfl;
    double rho = 1.0 / *parm_infl;
    for (size_t i = 0; i < k; ++i) w1[i + i * k] += (double)(ne - 1) * rho;
    /* :142 */
    int nrank;
    if (orc_mtx_eigen(ne, w1, el, ev, &nrank) != 0) {
        rc = 2;
        goto out;
    }
    /* :146-152 Pa = (V / lambda) V^T ; DGEMM('n','t') loop order J, L, I */
    for (size_t j = 0; j < k; ++j)
        for (size_t i = 0; i < k; ++i) w1[i + j * k] = ev[i + j * k] / el[j];
    for (size_t j = 0; j < k; ++j) {
        for (size_t i = 0; i < k; ++i) pa[i + j * k] = 0.0;
        for (size_t l = 0; l < k; ++l) {
            const double t = ev[j + l * k];
            for (size_t i = 0; i < k; ++i) pa[i + j * k] += t * w1[i + l * k];
        }
    }
    /* :164-165 work2 = Pa yr^T (k x p) */
    for (size_t j = 0; j < p; ++j) {
        for (size_t i = 0; i < k; ++i) w2[i + j * k] = 0.0;
        for (size_t l = 0; l < k; ++l) {
            const double t = yr[j + l * p];
            for (size_t i = 0; i < k; ++i) w2[i + j * k] += t * pa[i + l * k];
        }
    }
    /* :177-182 */
    for (size_t i = 0; i < k; ++i) {
        double t = w2[i] * dep[0];
        for (size_t j = 1; j < p; ++j) t += w2[i + j * k] * dep[j];
        w3[i] = t;
    }
    /* :186-193 T = V sqrt((k-1)/lambda) V^T */
    for (size_t j = 0; j < k; ++j) {
        const double s = sqrt((double)(ne - 1) / el[j]);
        for (size_t i = 0; i < k; ++i) w1[i + j * k] = ev[i + j * k] * s;
    }
    for (size_t j = 0; j < k; ++j) {
        for (size_t i = 0; i < k; ++i) trans[i + j * k] = 0.0;
        for (size_t l = 0; l < k; ++l) {
            const double t = ev[j + l * k];
            for (size_t i = 0; i < k; ++i) trans[i + j * k] += t * w1[i + l * k];
        }
    }
    /* :205-214 */
    if (transm) {
        memcpy(transm, w3, sizeof(double) * k);
    } else {
        for (size_t j = 0; j < k; ++j)
            for (size_t i = 0; i < k; ++i) trans[i + j * k] += w3[i];
    }
    if (pao) memcpy(pao, pa, sizeof(double) * k * k);
    /* :218-239 adaptive inflation */
    {
        double p1 = 0.0, p2 = 0.0, p3 = 0.0;
        for (size_t i = 0; i < p; ++i) {
            const double v = dep[i] * dep[i] / rdiag[i];
            p1 += rdiag_wloc ? v : v * rloc[i];
        }
        for (size_t j = 0; j < k; ++j)
            for (size_t i = 0; i < p; ++i) p2 += yr[i + j * p] * hdxb[i + j * (size_t)ldh];
        p2 /= (double)(ne - 1);
        for (size_t i = 0; i < p; ++i) p3 += rloc[i];
        const double p4 = (p1 - p3) / p2 - *parm_infl;
        const double q = (*parm_infl * p2 + p3) / p2;
        const double sigma_o = 2.0 / p3 * (q * q);
        const double sb2 = 0.04 * 0.04;
        const double gain = sb2 / (sigma_o + sb2);
        *parm_infl = *parm_infl + gain * p4;
    }
out:
    free(yr);
    free(w1);
    free(ev);
    free(el);
    free(pa);
    free(w2);
    free(w3);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Local observation selection                                               */
/* ------------------------------------------------------------------------- */

/* great-circle distance, acos form: common/common.f90:401-424 */
double orc_distll_1(double alon, double alat, double blon, double blat)
{
    const double r180 = 1.0 / 180.0;
    const double lon1 = alon * ORC_PI * r180;
    const double lon2 = blon * ORC_PI * r180;
    const double lat1 = alat * ORC_PI * r180;
    const double lat2 = blat * ORC_PI * r180;
    double cosd = sin(lat1) * sin(lat2) + cos(lat1) * cos(lat2) * cos(lon2 - lon1);
    cosd = cosd < 1.0 ? cosd : 1.0;
    cosd = cosd > -1.0 ? cosd : -1.0;
    return acos(cosd) * ORC_RE;
}

/*
 * Lorenz-96 inline selection, lorenz96/letkf/letkf.f90:216-254 (nwindow=1,
 * msw_local=0).  obs_gidx[i] is the grid index (0-based) carrying the maximum
 * of H row i (:226-228).  Keeps dist < xmaxloc and rho >= 1e-4.
 * Writes the selected obs indices and rho; returns the count.
 */
int orc_obs_local_l96(int nx, int ny, const int *obs_gidx, int ix, double xlocal, int *idx_out,
                      double *rloc_out)
{
    const double xmaxloc = xlocal * 2.0 * sqrt(10.0 / 3.0);
    int n = 0;
    for (int i = 0; i < ny; ++i) {
        int dj = abs(obs_gidx[i] - ix);
        int dm = nx - dj < dj ? nx - dj : dj;
        double dist = (double)dm;
        if (dist < xmaxloc) {
            /* the reference evaluates -0.5 * (dist/xlocal)**2 with a default-real 0.5
               literal; -0.5 is exact in binary so the product is the same */
            double r = exp(-0.5 * (dist / xlocal) * (dist / xlocal)) * exp(-0.5 * 0.0);
            if (r < 0.0001) continue;
            idx_out[n] = i;
            rloc_out[n] = r;
            ++n;
        }
    }
    return n;
}

/*
 * Global lat/lon bin index: speedy/letkf/letkf_obs.f90:243-313.
 * Stable sort of obs into lat bands j (lat[j] <= phi < lat[j+1], j = 0..nlat-2)
 * and then lon cells i (lon[i] <= lambda < lon[i+1], last cell up to 360).
 * Obs outside every band (poleward of lat[0] / lat[nlat-1]) are dropped.
 * perm[m] = input index of the m-th sorted obs; nobsgrd is nlon x nlat
 * (column-major, cumulative end offsets, band nlat-1 left at 0 as in the
 * reference).  Returns the number of binned obs.
 */
int orc_obs_bins_global(int nobs, const double *olon, const double *olat, int nlon, int nlat,
                        const double *lon, const double *lat, int *perm, int *nobsgrd)
{
    memset(nobsgrd, 0, sizeof(int) * (size_t)nlon * nlat);
    int *tmp = (int *)malloc(sizeof(int) * (nobs > 0 ? nobs : 1));
    int base = 0;
    for (int j = 0; j < nlat - 1; ++j) {
        int nj = 0;
        for (int n = 0; n < nobs; ++n) {
            if (olat[n] < lat[j] || lat[j + 1] <= olat[n]) continue;
            tmp[nj++] = n;
        }
        if (nj == 0) {
            for (int i = 0; i < nlon; ++i) nobsgrd[i + j * nlon] = base;
            continue;
        }
        int nn = 0;
        for (int i = 0; i < nlon; ++i) {
            for (int q = 0; q < nj; ++q) {
                const int n = tmp[q];
                if (i < nlon - 1) {
                    if (olon[n] < lon[i] || lon[i + 1] <= olon[n]) continue;
                } else {
                    if (olon[n] < lon[nlon - 1] || 360.0 <= olon[n]) continue;
                }
                perm[base + nn++] = n;
            }
            nobsgrd[i + j * nlon] = base + nn;
        }
        base += nn;
    }
    free(tmp);
    return base;
}

/* CSR ranges of the bin index for a lon range: speedy/letkf/letkf_tools.f90:575-603 */
static int bins_sub(int imin, int imax, int jmin, int jmax, int nlon, const int *nobsgrd,
                    int nn, int *use)
{
    for (int j = jmin; j <= jmax; ++j) { /* 1-based j, i as in the reference */
        int ib;
        if (imin > 1)
            ib = nobsgrd[(imin - 2) + (j - 1) * nlon] + 1;
        else
            ib = (j > 1) ? nobsgrd[(nlon - 1) + (j - 2) * nlon] + 1 : 1;
        const int ie = nobsgrd[(imax - 1) + (j - 1) * nlon];
        for (int ip = ib; ip <= ie; ++ip) use[nn++] = ip - 1;
    }
    return nn;
}

static int first_lt(double x, const double *tab, int nmax)
{
    /* DO j=1,nmax: IF (x < tab(j+1)) EXIT -> j in 1..nmax+1 (1-based) */
    int j = 1;
    for (; j <= nmax; ++j)
        if (x < tab[j]) break;
    return j;
}

/*
 * Candidate obs (sorted-array indices) in the lat/lon rectangle around a grid
 * point, with the 0/360 wrap cases: speedy/letkf/letkf_tools.f90:302-451.
 * Returns the count; order is the reference's scan order.
 */
int orc_rect_candidates(double lon1, double lat1, double dlon_zero, double dlat_zero, int nlon,
                        int nlat, const double *lon, const double *lat, const int *nobsgrd,
                        int *use)
{
    double minlon = lon1 - dlon_zero, maxlon = lon1 + dlon_zero;
    const double minlat = lat1 - dlat_zero, maxlat = lat1 + dlat_zero;
    if (maxlon - minlon >= 360.0) {
        minlon = 0.0;
        maxlon = 360.0;
    }
    const int jmin = first_lt(minlat, lat, nlat - 2);
    const int jmax = first_lt(maxlat, lat, nlat - 2);
    int nn = 0, imin, imax;
    if (minlon >= 0 && maxlon <= 360.0) {
        imin = first_lt(minlon, lon, nlon - 1);
        imax = first_lt(maxlon, lon, nlon - 1);
        nn = bins_sub(imin, imax, jmin, jmax, nlon, nobsgrd, nn, use);
    } else if (minlon >= 0 && maxlon > 360.0) {
        imin = first_lt(minlon, lon, nlon - 1);
        maxlon -= 360.0;
        if (maxlon > 360.0) {
            nn = bins_sub(1, nlon, jmin, jmax, nlon, nobsgrd, nn, use);
        } else {
            imax = first_lt(maxlon, lon, nlon - 1);
            if (imax > imin) {
                nn = bins_sub(1, nlon, jmin, jmax, nlon, nobsgrd, nn, use);
            } else {
                nn = bins_sub(1, imax, jmin, jmax, nlon, nobsgrd, nn, use);
                imin = first_lt(minlon, lon, nlon - 1);
                nn = bins_sub(imin, nlon, jmin, jmax, nlon, nobsgrd, nn, use);
            }
        }
    } else if (minlon < 0 && maxlon <= 360.0) {
        imax = first_lt(maxlon, lon, nlon - 1);
        minlon += 360.0;
        if (minlon < 0) {
            nn = bins_sub(1, nlon, jmin, jmax, nlon, nobsgrd, nn, use);
        } else {
            imin = first_lt(minlon, lon, nlon - 1);
            if (imin < imax) {
                nn = bins_sub(1, nlon, jmin, jmax, nlon, nobsgrd, nn, use);
            } else {
                nn = bins_sub(1, imax, jmin, jmax, nlon, nobsgrd, nn, use);
                imin = first_lt(minlon, lon, nlon - 1);
                nn = bins_sub(imin, nlon, jmin, jmax, nlon, nobsgrd, nn, use);
            }
        }
    } else {
        maxlon -= 360.0;
        minlon += 360.0;
        if (maxlon > 360.0 || minlon < 0) {
            nn = bins_sub(1, nlon, jmin, jmax, nlon, nobsgrd, nn, use);
        } else {
            imin = first_lt(minlon, lon, nlon - 1);
            imax = first_lt(maxlon, lon, nlon - 1);
            if (imin > imax)
                nn = bins_sub(1, nlon, jmin, jmax, nlon, nobsgrd, nn, use);
            else
                nn = bins_sub(imin, imax, jmin, jmax, nlon, nobsgrd, nn, use);
        }
    }
    return nn;
}

static int iobs_of(int elm)
{
    switch (elm) {
    case ID_U: return 0;
    case ID_V: return 1;
    case ID_T: return 2;
    case ID_Q: return 3;
    case ID_RH: return 4;
    case ID_PS: return 5;
    case ID_RAIN: return 6;
    default: return -1;
    }
}

/*
 * obs_local for a global grid: speedy/letkf/letkf_tools.f90:272-573.
 *   elm/olon/olat/olev/odat/oerr: sorted obs arrays (obs_bins order)
 *   ilev: 1-based model level (1 => surface-pressure obs get dlev = 0)
 *   logpfm: ln of the background-mean full pressure at the point
 *   var_local_row[7]: localisation factor per obs type (u,v,t,q,rh,ps,rain)
 * Writes selected sorted-array indices, rloc and rdiag=err^2; returns nobsl.
 * `use` is caller scratch of at least nobs ints.
 */
int orc_obs_local_global(double lon1, double lat1, double dlon_zero, double dlat_zero,
                         double logpfm, int ilev, int nlon, int nlat, const double *lon,
                         const double *lat, const int *nobsgrd, const double *elm,
                         const double *olon, const double *olat, const double *olev,
                         const double *odat, const double *oerr, double dist_zero,
                         double dist_zerov, double sigma_obs, double sigma_obsv,
                         const double *var_local_row, int *use, int *idx_out, double *rloc_out,
                         double *rdiag_out)
{
    const int nn = orc_rect_candidates(lon1, lat1, dlon_zero, dlat_zero, nlon, nlat, lon, lat,
                                       nobsgrd, use);
    int nobsl = 0;
    for (int q = 0; q < nn; ++q) {
        const int o = use[q];
        const int e = (int)lround(elm[o]);
        double dlev;
        if (e == ID_PS && ilev > 1)
            dlev = fabs(log(odat[o]) - logpfm);
        else if (e != ID_PS)
            dlev = fabs(log(olev[o]) - logpfm);
        else
            dlev = 0.0;
        if (dlev > dist_zerov) continue;
        const double dist = orc_distll_1(olon[o], olat[o], lon1, lat1);
        if (dist > dist_zero) continue;
        const int io = iobs_of(e);
        const double vl = io >= 0 ? var_local_row[io] : 0.0;
        if (vl < DBL_MIN) continue;
        idx_out[nobsl] = o;
        rdiag_out[nobsl] = oerr[o] * oerr[o];
        rloc_out[nobsl] = exp(-0.5 * ((dist / sigma_obs) * (dist / sigma_obs) +
                                      (dlev / sigma_obsv) * (dlev / sigma_obsv))) *
                          vl;
        ++nobsl;
    }
    return nobsl;
}

/* ------------------------------------------------------------------------- */
/* Regional selection with the MAX_NOBS_PER_GRID cap                          */
/* scale/letkf/letkf_tools.f90:853-1178                                       */
/* ------------------------------------------------------------------------- */

/*
 * Regional obs index (scale/letkf/letkf_obs.f90:740-830, one subdomain):
 * obs n with qc good goes to cell i = ceiling(ri-0.5), j = ceiling(rj-0.5)
 * (1-based grid units), clamped to [1,nx] x [1,ny]; obs keep their input
 * order inside a cell.  nobsgrd is (0:nx) x ny column-major, cumulative,
 * with nobsgrd(0,j) = nobsgrd(nx,j-1).  perm[m] = input index of sorted obs
 * m.  Returns the number of binned obs.
 */
int orc_obs_bins_regional(int nobs, const double *ri, const double *rj, const int *qc, int nx,
                          int ny, int *perm, int *nobsgrd)
{
    const int ld = nx + 1;
    int *cell = (int *)malloc(sizeof(int) * (nobs > 0 ? nobs : 1));
    int *cnt = (int *)calloc((size_t)nx * ny, sizeof(int));
    for (int n = 0; n < nobs; ++n) {
        cell[n] = -1;
        if (qc && qc[n] != 1) continue;
        int i = (int)ceil(ri[n] - 0.5), j = (int)ceil(rj[n] - 0.5);
        if (i < 1) i = 1;
        if (i > nx) i = nx;
        if (j < 1) j = 1;
        if (j > ny) j = ny;
        cell[n] = (i - 1) + nx * (j - 1);
        cnt[cell[n]]++;
    }
    memset(nobsgrd, 0, sizeof(int) * (size_t)ld * ny);
    for (int j = 0; j < ny; ++j) {
        if (j > 0) nobsgrd[0 + ld * j] = nobsgrd[nx + ld * (j - 1)];
        for (int i = 1; i <= nx; ++i)
            nobsgrd[i + ld * j] = nobsgrd[(i - 1) + ld * j] + cnt[(i - 1) + nx * j];
    }
    int *next = (int *)malloc(sizeof(int) * (size_t)nx * ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 1; i <= nx; ++i) next[(i - 1) + nx * j] = nobsgrd[(i - 1) + ld * j];
    int tot = 0;
    for (int n = 0; n < nobs; ++n) {
        if (cell[n] < 0) continue;
        perm[next[cell[n]]++] = n;
        ++tot;
    }
    free(cell);
    free(cnt);
    free(next);
    return tot;
}

/*
 * Regional obs_local (scale/letkf/letkf_tools.f90:853-1178, one subdomain, radar-type
 * vertical distance):
 *   rectangle imin = max(1, floor(ri - dlon_zero)), imax = min(nx, ceiling(ri + dlon_zero)),
 *   likewise j (:906-909); candidates scanned row by row through nobsgrd (obs_choose,
 *   scale/letkf/letkf_obs.f90:1054-1119);
 *   rdx = (ri - ri_o) DX, rdy = (rj - rj_o) DY, nd_h = sqrt(rdx^2 + rdy^2) / sigma_h,
 *   nd_v = |z_o - rz| / sigma_v (:972-1004), ndist = nd_h^2 + nd_v^2 rejected when above
 *   the literal 13.33333333 (scale/letkf/letkf_obs.f90:29-30; :1022-1023);
 *   rloc = exp(-0.5 ndist), rdiag = err^2 / rloc (:1027-1037);
 *   nobs_max > 0: per bucket (the reference's (element, obs type) pair) keep the
 *   nobs_max smallest rdiag with the reference's slot/sort structure (:1041-1128) and
 *   emit bucket by bucket in slot order (:1134-1156); nobs_max <= 0: scan order.
 * Scratch: isort, slot_obs (nbkt*nobs_max ints), slot_rd, slot_rl (doubles), cnt (nbkt).
 * Returns nobsl.
 */
int orc_obs_local_regional(double ri, double rj, double rz, int nx, int ny, const int *nobsgrd,
                           const double *ori, const double *orj, const double *oz,
                           const double *oerr, const int *obkt, int nbkt, double dx, double dy,
                           double sigma_h, double sigma_v, double dlon_zero, double dlat_zero,
                           int nobs_max, int *isort, int *slot_obs, double *slot_rd,
                           double *slot_rl, int *cnt, int *idx_out, double *rloc_out,
                           double *rdiag_out)
{
    const double nd_max = 13.33333333;
    const int ld = nx + 1;
    int imin = (int)floor(ri - dlon_zero), imax = (int)ceil(ri + dlon_zero);
    int jmin = (int)floor(rj - dlat_zero), jmax = (int)ceil(rj + dlat_zero);
    if (imin < 1) imin = 1;
    if (imax > nx) imax = nx;
    if (jmin < 1) jmin = 1;
    if (jmax > ny) jmax = ny;
    int nsel = 0;
    const int cap = nobs_max;
    if (cap > 0) {
        for (int b = 0; b < nbkt; ++b) cnt[b] = 0;
        for (size_t t = 0; t < (size_t)nbkt * cap; ++t) isort[t] = -1;
    }
    if (imin > imax || jmin > jmax) return 0;
    for (int j = jmin; j <= jmax; ++j) {
        const int beg = nobsgrd[(imin - 1) + ld * (j - 1)];
        const int end = nobsgrd[imax + ld * (j - 1)];
        for (int o = beg; o < end; ++o) {
            const double rdx = (ri - ori[o]) * dx;
            const double rdy = (rj - orj[o]) * dy;
            const double nd_h = sqrt(rdx * rdx + rdy * rdy) / sigma_h;
            const double nd_v = fabs(oz[o] - rz) / sigma_v;
            const double ndist = nd_h * nd_h + nd_v * nd_v;
            if (ndist > nd_max) continue;
            const double nrloc = exp(-0.5 * ndist);
            const double nrdiag = oerr[o] * oerr[o] / nrloc;
            if (cap <= 0) {
                idx_out[nsel] = o;
                rloc_out[nsel] = nrloc;
                rdiag_out[nsel] = nrdiag;
                ++nsel;
                continue;
            }
            const int b = obkt[o];
            int *is = isort + (size_t)b * cap;
            int *so = slot_obs + (size_t)b * cap;
            double *sr = slot_rd + (size_t)b * cap;
            double *sl = slot_rl + (size_t)b * cap;
            /* case 0: full and not better than the last-priority obs */
            if (cnt[b] >= cap && nrdiag >= sr[is[cap - 1]]) continue;
            for (int s2 = 0; s2 < cap; ++s2) {
                if (is[s2] < 0) { /* case 1: spare slot at the end */
                    const int slot = cnt[b]++;
                    is[s2] = slot;
                    so[slot] = o;
                    sr[slot] = nrdiag;
                    sl[slot] = nrloc;
                    break;
                } else if (nrdiag < sr[is[s2]]) {
                    int slot;
                    if (cnt[b] < cap) { /* case 2: insert, spare slot */
                        slot = cnt[b]++;
                        for (int ss = cnt[b] - 1; ss > s2; --ss) is[ss] = is[ss - 1];
                    } else { /* case 3: overwrite the last-priority slot */
                        slot = is[cap - 1];
                        for (int ss = cap - 1; ss > s2; --ss) is[ss] = is[ss - 1];
                    }
                    is[s2] = slot;
                    so[slot] = o;
                    sr[slot] = nrdiag;
                    sl[slot] = nrloc;
                    break;
                }
            }
        }
    }
    if (cap <= 0) return nsel;
    for (int b = 0; b < nbkt; ++b)
        for (int s2 = 0; s2 < cnt[b]; ++s2) {
            idx_out[nsel] = slot_obs[(size_t)b * cap + s2];
            rdiag_out[nsel] = slot_rd[(size_t)b * cap + s2];
            rloc_out[nsel] = slot_rl[(size_t)b * cap + s2];
            ++nsel;
        }
    return nsel;
}

/*
 * The reference's cap over one candidate stream (scale/letkf/letkf_tools.f90:
 * 1041-1156, one (element, obs type) bucket): candidates 0..n-1 in scan order
 * with priority key[i] (smaller first).  Case 0 skips a candidate when the set
 * is full and key >= the last-priority key; otherwise the first sorted
 * position s with key < key(s) (strict, so an earlier candidate wins a tie),
 * or the first empty one, takes it: into the next storage slot while the set
 * is not full (cases 1, 2), else into the storage slot of the last-priority
 * entry, which it evicts (case 3).  slot_src[slot] = candidate in that slot,
 * the reference's output order (:1139-1150).  The global grid's nearest-first
 * cap uses key = -rho.  Scratch isort: cap ints.  Returns the kept count.
 */
int orc_cap_stream(int n, const double *key, int cap, int *isort, int *slot_src)
{
    int cnt = 0;
    for (int s2 = 0; s2 < cap; ++s2) isort[s2] = -1;
    for (int i = 0; i < n; ++i) {
        const double k = key[i];
        if (cnt >= cap && k >= key[slot_src[isort[cap - 1]]]) continue; /* case 0 */
        for (int s2 = 0; s2 < cap; ++s2) {
            if (isort[s2] < 0) { /* case 1 */
                const int slot = cnt++;
                isort[s2] = slot;
                slot_src[slot] = i;
                break;
            } else if (k < key[slot_src[isort[s2]]]) {
                int slot;
                if (cnt < cap) { /* case 2 */
                    slot = cnt++;
                    for (int ss = cnt - 1; ss > s2; --ss) isort[ss] = isort[ss - 1];
                } else { /* case 3 */
                    slot = isort[cap - 1];
                    for (int ss = cap - 1; ss > s2; --ss) isort[ss] = isort[ss - 1];
                }
                isort[s2] = slot;
                slot_src[slot] = i;
                break;
            }
        }
    }
    return cnt;
}

/* ------------------------------------------------------------------------- */
/* Driver: das_letkf for a global sigma-level grid                            */
/* speedy/letkf/letkf_tools.f90:45-267                                        */
/* ------------------------------------------------------------------------- */

/*
 * Global das_letkf restatement (one rank owning every grid point).
 *   gues3d  : nij x nlev x ne x nv3d (column-major, ij fastest); overwritten
 *             with perturbations as in the reference (:95-112)
 *   gues2d  : nij x ne x nv2d
 *   anal3d / anal2d : outputs, same shapes
 *   lon1/lat1[nij] grid point coordinates; sig[nlev]; ips: index of ps among
 *   the 2-D variables (ln p_full uses ps_mean * sig)
 *   infl3d  : nij x nlev (one inflation field shared by the variables, which
 *             all share one var_local row); in/out.  infl2d: nij x nv2d.
 *   obs arrays: sorted (obs_bins order); hdxf is nobs x ne column-major.
 *   q_top_lev: levels >= q_top_lev (1-based) keep the background q (:172-177);
 *   iq: index of q among the 3-D variables (<0: rule off).
 *   nobsl_out[nij*nlev] optional.
 * Returns 0 or the first nonzero letkf_core status.
 */
int orc_das_global(int nij, int nlev, int ne, int nv3d, int nv2d, double *gues3d, double *gues2d,
                   double *anal3d, double *anal2d, const double *lon1, const double *lat1,
                   const double *sig, int ips, double *infl3d, double *infl2d, int nlon, int nlat,
                   const double *lon, const double *lat, int nobs, const int *nobsgrd,
                   const double *elm, const double *olon, const double *olat, const double *olev,
                   const double *odat, const double *oerr, const double *odep,
                   const double *hdxf, double sigma_obs, double sigma_obsv,
                   const double *var_local_row, int q_top_lev, int iq, int *nobsl_out)
{
    const size_t NIJ = (size_t)nij, NL = (size_t)nlev, K = (size_t)ne;
    if (nobs == 0) { /* :73-78 no observation: analysis = background */
        memcpy(anal3d, gues3d, sizeof(double) * NIJ * NL * K * (size_t)nv3d);
        memcpy(anal2d, gues2d, sizeof(double) * NIJ * K * (size_t)nv2d);
        return 0;
    }
#define G3(i, l, m, n) gues3d[(i) + NIJ * ((l) + NL * ((m) + K * (size_t)(n)))]
#define A3(i, l, m, n) anal3d[(i) + NIJ * ((l) + NL * ((m) + K * (size_t)(n)))]
#define G2(i, m, n) gues2d[(i) + NIJ * ((m) + K * (size_t)(n))]
#define A2(i, m, n) anal2d[(i) + NIJ * ((m) + K * (size_t)(n))]
    double *mean3d = (double *)malloc(sizeof(double) * NIJ * NL * (nv3d > 0 ? nv3d : 1));
    double *mean2d = (double *)malloc(sizeof(double) * NIJ * (nv2d > 0 ? nv2d : 1));
    /* ensmean_grd: speedy/common/common_speedy.f90:335-370 */
    for (int n = 0; n < nv3d; ++n)
        for (size_t l = 0; l < NL; ++l)
            for (size_t i = 0; i < NIJ; ++i) {
                double s = G3(i, l, 0, n);
                for (size_t m = 1; m < K; ++m) s += G3(i, l, m, n);
                mean3d[i + NIJ * (l + NL * n)] = s / (double)ne;
            }
    for (int n = 0; n < nv2d; ++n)
        for (size_t i = 0; i < NIJ; ++i) {
            double s = G2(i, 0, n);
            for (size_t m = 1; m < K; ++m) s += G2(i, m, n);
            mean2d[i + NIJ * n] = s / (double)ne;
        }
    for (int n = 0; n < nv3d; ++n)
        for (size_t m = 0; m < K; ++m)
            for (size_t l = 0; l < NL; ++l)
                for (size_t i = 0; i < NIJ; ++i) G3(i, l, m, n) -= mean3d[i + NIJ * (l + NL * n)];
    for (int n = 0; n < nv2d; ++n)
        for (size_t m = 0; m < K; ++m)
            for (size_t i = 0; i < NIJ; ++i) G2(i, m, n) -= mean2d[i + NIJ * n];
    /* localisation constants: speedy/letkf/letkf_obs.f90:82-88 */
    const double dist_zero = sigma_obs * sqrt(10.0 / 3.0) * 2.0;
    const double dist_zerov = sigma_obsv * sqrt(10.0 / 3.0) * 2.0;
    const double dlat_zero = dist_zero / ORC_PI / ORC_RE * 180.0;
    int *use = (int *)malloc(sizeof(int) * (2 * (size_t)nobs + 1));
    int *idx = (int *)malloc(sizeof(int) * (2 * (size_t)nobs + 1));
    double *rl = (double *)malloc(sizeof(double) * (2 * (size_t)nobs + 1));
    double *rd = (double *)malloc(sizeof(double) * (2 * (size_t)nobs + 1));
    double *dp = (double *)malloc(sizeof(double) * (2 * (size_t)nobs + 1));
    double *hl = (double *)malloc(sizeof(double) * (2 * (size_t)nobs + 1) * K);
    double *tr = (double *)malloc(sizeof(double) * K * K);
    int rc = 0;
    for (size_t l = 0; l < NL && rc == 0; ++l) {
        for (size_t ij = 0; ij < NIJ; ++ij) {
            const double dlon_zero = dlat_zero / cos(ORC_PI * lat1[ij] / 180.0);
            const double logpfm = log(mean2d[ij + NIJ * ips] * sig[l]);
            int nobsl = 0;
            if (nobs > 0)
                nobsl = orc_obs_local_global(lon1[ij], lat1[ij], dlon_zero, dlat_zero, logpfm,
                                             (int)l + 1, nlon, nlat, lon, lat, nobsgrd, elm, olon,
                                             olat, olev, odat, oerr, dist_zero, dist_zerov,
                                             sigma_obs, sigma_obsv, var_local_row, use, idx, rl,
                                             rd);
            if (nobsl_out) nobsl_out[ij + NIJ * l] = nobsl;
            for (int q = 0; q < nobsl; ++q) {
                dp[q] = odep[idx[q]];
                for (size_t m = 0; m < K; ++m) hl[q + (size_t)nobsl * m] = hdxf[idx[q] + (size_t)nobs * m];
            }
            double parm = infl3d[ij + NIJ * l];
            int st = orc_letkf_core(ne, nobsl > 0 ? nobsl : 1, nobsl, hl, rd, rl, dp, &parm, tr,
                                    NULL, NULL, 0, NULL);
            if (st) {
                rc = st;
                break;
            }
            infl3d[ij + NIJ * l] = parm;
            for (int n = 0; n < nv3d; ++n) {
                const int frozen = (n == iq && (int)l + 1 >= q_top_lev);
                for (size_t m = 0; m < K; ++m) {
                    double a = mean3d[ij + NIJ * (l + NL * n)];
                    if (frozen) {
                        a += G3(ij, l, m, n);
                    } else {
                        for (size_t kk = 0; kk < K; ++kk) a += G3(ij, l, kk, n) * tr[kk + K * m];
                    }
                    A3(ij, l, m, n) = a;
                }
            }
            if (l == 0) {
                for (int n = 0; n < nv2d; ++n) {
                    /* every 2-D variable reuses the level-1 transform; the reference copies
                       the inflation from work2d(ij, n2n) with a 3-D index (:178-183), so the
                       2-D field takes column 0 of itself and is never updated */
                    infl2d[ij + NIJ * n] = infl2d[ij];
                    for (size_t m = 0; m < K; ++m) {
                        double a = mean2d[ij + NIJ * n];
                        for (size_t kk = 0; kk < K; ++kk) a += G2(ij, kk, n) * tr[kk + K * m];
                        A2(ij, m, n) = a;
                    }
                }
            }
        }
    }
    free(mean3d);
    free(mean2d);
    free(use);
    free(idx);
    free(rl);
    free(rd);
    free(dp);
    free(hl);
    free(tr);
    return rc;
#undef G3
#undef A3
#undef G2
#undef A2
}

/* ------------------------------------------------------------------------- */
/* Obs-operator indexing: speedy/common/common_obs_speedy.f90:137-282          */
/* ------------------------------------------------------------------------- */

/*
 * lon -> ri and lat -> rj of phys2ijk (:156-175).  Returns 0 when inside the
 * grid, 1 when the reference's range checks (speedy/letkf/obsope.f90:57-66)
 * reject the obs.
 */
int orc_phys2ij(double rlon, double rlat, int nlon, int nlat, const double *lat, double *ri,
                double *rj)
{
    if (rlon == 0.0 || rlon == 360.0)
        *ri = (double)(nlon + 1);
    else
        *ri = rlon / 360.0 * (double)nlon + 1.0;
    if (ceil(*ri) < 2 || nlon + 1 < ceil(*ri)) return 1;
    int j = 1;
    for (; j <= nlat; ++j)
        if (rlat < lat[j - 1]) break;
    if (j == 1) {
        *rj = (rlat + 90.0) / (lat[0] + 90.0);
    } else if (j == nlat + 1) {
        const double aj = (rlat - lat[nlat - 1]) / (90.0 - lat[nlat - 1]);
        *rj = (double)nlat + aj;
    } else {
        const double aj = (rlat - lat[j - 2]) / (lat[j - 1] - lat[j - 2]);
        *rj = (double)(j - 1) + aj;
    }
    if (ceil(*rj) < 2 || nlat < ceil(*rj)) return 1;
    return 0;
}

/* bilinear interpolation with lon wrap: common_obs_speedy.f90:215-242 (var nlon x nlat) */
double orc_itpl_2d(const double *var, int nlon, double ri, double rj)
{
    const int i = (int)ceil(ri);
    const double ai = ri - (double)(i - 1);
    const int j = (int)ceil(rj);
    const double aj = rj - (double)(j - 1);
#define V(ii, jj) var[((ii)-1) + (size_t)nlon * ((jj)-1)]
    const int ie = i <= nlon ? i : 1;
    return V(i - 1, j - 1) * (1 - ai) * (1 - aj) + V(ie, j - 1) * ai * (1 - aj) +
           V(i - 1, j) * (1 - ai) * aj + V(ie, j) * ai * aj;
#undef V
}

/* trilinear interpolation with lon wrap: common_obs_speedy.f90:244-282 (nlon x nlat x nlev) */
double orc_itpl_3d(const double *var, int nlon, int nlat, double ri, double rj, double rk)
{
    const int i = (int)ceil(ri);
    const double ai = ri - (double)(i - 1);
    const int j = (int)ceil(rj);
    const double aj = rj - (double)(j - 1);
    const int k = (int)ceil(rk);
    const double ak = rk - (double)(k - 1);
#define V(ii, jj, kk) var[((ii)-1) + (size_t)nlon * (((jj)-1) + (size_t)nlat * ((kk)-1))]
    const int ie = i <= nlon ? i : 1;
    return V(i - 1, j - 1, k - 1) * (1 - ai) * (1 - aj) * (1 - ak) +
           V(ie, j - 1, k - 1) * ai * (1 - aj) * (1 - ak) +
           V(i - 1, j, k - 1) * (1 - ai) * aj * (1 - ak) + V(ie, j, k - 1) * ai * aj * (1 - ak) +
           V(i - 1, j - 1, k) * (1 - ai) * (1 - aj) * ak + V(ie, j - 1, k) * ai * (1 - aj) * ak +
           V(i - 1, j, k) * (1 - ai) * aj * ak + V(ie, j, k) * ai * aj * ak;
#undef V
}

/*
 * Full obs operator for one member (Trans_XtoY + phys2ijk + the obsope range
 * rules): common_obs_speedy.f90:30-81,137-211; speedy/letkf/obsope.f90:56-89.
 * v3d: nlon x nlat x nlev x nv3d (u,v,t,q), ps: nlon x nlat, sig[nlev].
 * Returns 0 with *yobs set, or 1 when the obs is rejected (qc = 0).
 * rk_out / ri_out / rj_out receive the fractional grid coordinates.
 */
int orc_trans_xtoy(int elm, double rlon, double rlat, double rlev, double odat, int nlon,
                   int nlat, int nlev, const double *lat, const double *sig, const double *v3d,
                   const double *ps, double *yobs, double *ri_out, double *rj_out, double *rk_out)
{
    double ri, rj, rk;
    if (orc_phys2ij(rlon, rlat, nlon, nlat, lat, &ri, &rj)) return 1;
    if (elm == ID_PS) {
        rk = 0.0;
    } else {
        /* ln p_full interpolated at the obs column for every level (:177-190) */
        double plev[256];
        const int i = (int)ceil(ri), j = (int)ceil(rj);
        double lnps[4];
        for (int k = 0; k < nlev; ++k) {
            const int ie = i <= nlon ? i : 1;
            lnps[0] = log(ps[(i - 2) + (size_t)nlon * (j - 2)] * sig[k]);
            lnps[1] = log(ps[(ie - 1) + (size_t)nlon * (j - 2)] * sig[k]);
            lnps[2] = log(ps[(i - 2) + (size_t)nlon * (j - 1)] * sig[k]);
            lnps[3] = log(ps[(ie - 1) + (size_t)nlon * (j - 1)] * sig[k]);
            const double ai = ri - (double)(i - 1), aj = rj - (double)(j - 1);
            plev[k] = lnps[0] * (1 - ai) * (1 - aj) + lnps[1] * ai * (1 - aj) +
                      lnps[2] * (1 - ai) * aj + lnps[3] * ai * aj;
        }
        rk = log(rlev);
        int k = 2;
        for (; k <= nlev - 1; ++k)
            if (plev[k - 1] < rk) break;
        const double ak = (rk - plev[k - 2]) / (plev[k - 1] - plev[k - 2]);
        rk = (double)(k - 1) + ak;
    }
    /* obsope.f90:67-89 */
    if (ceil(rk) > nlev) return 1;
    if (ceil(rk) < 2 && elm != ID_PS) {
        if (elm == ID_U || elm == ID_V)
            rk = 1.00001;
        else
            return 1;
    }
    if (elm == ID_PS && odat < -100.0) return 1;
    const size_t plane = (size_t)nlon * nlat * nlev;
    switch (elm) {
    case ID_U: *yobs = orc_itpl_3d(v3d + 0 * plane, nlon, nlat, ri, rj, rk); break;
    case ID_V: *yobs = orc_itpl_3d(v3d + 1 * plane, nlon, nlat, ri, rj, rk); break;
    case ID_T: *yobs = orc_itpl_3d(v3d + 2 * plane, nlon, nlat, ri, rj, rk); break;
    case ID_Q: *yobs = orc_itpl_3d(v3d + 3 * plane, nlon, nlat, ri, rj, rk); break;
    case ID_PS: *yobs = orc_itpl_2d(ps, nlon, ri, rj); break;
    default: return 1;
    }
    if (ri_out) *ri_out = ri;
    if (rj_out) *rj_out = rj;
    if (rk_out) *rk_out = rk;
    return 0;
}

/*
 * Relaxation of one point's analysis, scale/letkf/letkf_tools.f90:215-268
 * (weight_RTPP :1224-1236, weight_RTPS :1240-1268, relax_beta's beta, and the
 * q spread limit).  xb: nrow background rows of k members (xb[r*k + m]);
 * trans = Wa (k x k, column-major, letkf_core with transm), transm = w-bar,
 * pa (k x k, column-major); mode 0 none, 1 RTPP, 2 RTPS; q_row = the row of
 * the 3-D q variable (< 0: none).  Writes xa[r*k + m]; infl_out[r] (nullable)
 * gets the RTPS inflation of each row.
 */
void orc_relax_point(int k, int nrow, const double *xb, const double *trans, const double *transm,
                     const double *pa, int mode, double alpha, double beta, double q_sprd_max,
                     int q_row, double *xa, double *infl_out)
{
    double *g = (double *)malloc(sizeof(double) * k);
    double *wr = (double *)malloc(sizeof(double) * k * k);
    double *tr = (double *)malloc(sizeof(double) * k * k);
    for (int r = 0; r < nrow; ++r) {
        const double *x = xb + (size_t)r * k;
        double *a = xa + (size_t)r * k;
        double mean = 0.0;
        for (int m = 0; m < k; ++m) mean += x[m];
        mean /= (double)k;
        for (int m = 0; m < k; ++m) g[m] = x[m] - mean;
        double infl = 1.0;
        if (beta == 0.0) {
            for (int m = 0; m < k; ++m) a[m] = mean + g[m];
        } else {
            if (mode == 1) {
                for (int i = 0; i < k * k; ++i) wr[i] = (1.0 - alpha) * trans[i];
                for (int m = 0; m < k; ++m) wr[m + m * k] += alpha;
            } else if (mode == 2) {
                double var_g = 0.0, var_a = 0.0;
                for (int m = 0; m < k; ++m) {
                    var_g += g[m] * g[m];
                    for (int kk = 0; kk < k; ++kk) var_a += g[kk] * pa[kk + m * k] * g[m];
                }
                if (var_g > 0.0 && var_a > 0.0) {
                    infl = alpha * sqrt(var_g / (var_a * (double)(k - 1))) - alpha + 1.0;
                    for (int i = 0; i < k * k; ++i) wr[i] = trans[i] * infl;
                } else {
                    for (int i = 0; i < k * k; ++i) wr[i] = trans[i];
                }
            } else {
                for (int i = 0; i < k * k; ++i) wr[i] = trans[i];
            }
            for (int m = 0; m < k; ++m) {
                for (int kk = 0; kk < k; ++kk) tr[kk + m * k] = (wr[kk + m * k] + transm[kk]) * beta;
                tr[m + m * k] += 1.0 - beta;
            }
            for (int m = 0; m < k; ++m) {
                double s = mean;
                for (int kk = 0; kk < k; ++kk) s += g[kk] * tr[kk + m * k];
                a[m] = s;
            }
        }
        if (infl_out) infl_out[r] = infl;
        if (q_sprd_max > 0.0 && r == q_row) {
            double qm = 0.0, qs = 0.0;
            for (int m = 0; m < k; ++m) qm += a[m];
            qm /= (double)k;
            for (int m = 0; m < k; ++m) {
                g[m] = a[m] - qm;
                qs += g[m] * g[m];
            }
            qs = sqrt(qs / (double)(k - 1)) / qm;
            if (qs > q_sprd_max)
                for (int m = 0; m < k; ++m) a[m] = qm + g[m] * q_sprd_max / qs;
        }
    }
    free(g);
    free(wr);
    free(tr);
}

/*
 * Lorenz-96 model, lorenz96/model/lorenz96.f90:93-128 (tinteg_rk4) and
 * :290-309 (lorenz96_core) for any nx: kt RK4 steps of x (nx values) in
 * place.  Operations in the reference's order.
 */
static void orc_l96_core(int nx, const double *x, double *q, double dt, double force)
{
    for (int i = 0; i < nx; ++i) {
        const int im1 = i == 0 ? nx - 1 : i - 1;
        const int im2 = i < 2 ? i + nx - 2 : i - 2;
        const int ip1 = i == nx - 1 ? 0 : i + 1;
        q[i] = dt * (x[im1] * (x[ip1] - x[im2]) - x[i] + force);
    }
}

void orc_l96_rk4(int nx, int kt, double dt, double force, double *x)
{
    double *xt = (double *)malloc(sizeof(double) * nx * 5);
    double *q1 = xt + nx, *q2 = q1 + nx, *q3 = q2 + nx, *q4 = q3 + nx;
    for (int s = 0; s < kt; ++s) {
        orc_l96_core(nx, x, q1, dt, force);
        for (int i = 0; i < nx; ++i) xt[i] = x[i] + 0.5 * q1[i];
        orc_l96_core(nx, xt, q2, dt, force);
        for (int i = 0; i < nx; ++i) xt[i] = x[i] + 0.5 * q2[i];
        orc_l96_core(nx, xt, q3, dt, force);
        for (int i = 0; i < nx; ++i) xt[i] = x[i] + q3[i];
        orc_l96_core(nx, xt, q4, dt, force);
        for (int i = 0; i < nx; ++i) x[i] = x[i] + (q1[i] + 2.0 * q2[i] + 2.0 * q3[i] + q4[i]) / 6.0;
    }
    free(xt);
}


/* ------------------------------------------------------------------------- */
/* SCALE radar obs operator                                                  */
/* ------------------------------------------------------------------------- */

/* scale/common/common_obs_scale.f90:71-72,141-151 */
#define ID_RADAR_REF 4001
#define ID_RADAR_VR 4002
#define IQC_GOOD 0
#define IQC_REF_LOW 11
#define IQC_RADAR_VHI 19
#define IQC_OUT_VHI 20
#define IQC_OUT_VLO 21
#define IQC_OUT_H 22
#define IQC_OTYPE 90
/* common/common.f90:29-30,38-40 */
#define ORC_GG 9.81
#define ORC_RD 287.05
#define ORC_UNDEF (-9.99e33)

/* Default-real (single precision) literals of calc_ref_vr, promoted to double
 * where they meet double operands, as the Fortran compiler does. */
#define F4(x) ((double)(float)(x))

/*
 * phys2ijkz (scale/common/common_obs_scale.f90 phys2ijkz) on a grid of nx x ny
 * columns and nlev levels without halo.  hgt(ij, lev) column-major with
 * ij = (i-1) + nx (j-1); ri, rj, rk are 1-based grid coordinates.  The
 * reference's halo lets it accept ri = 1 and a height equal to the top level's;
 * here the corners must lie inside the grid: 1 < ri <= nx, 1 < rj <= ny, and a
 * height equal to the top level's takes the top layer.  Returns the qc.
 */
int orc_phys2ijkz(const double *hgt, int nx, int ny, int nlev, double ri, double rj,
                  double rlev, double *rk)
{
    const long nij = (long)nx * ny;
    *rk = ORC_UNDEF;
    if (!(ri > 1.0 && ri <= (double)nx && rj > 1.0 && rj <= (double)ny)) return IQC_OUT_H;
    const int i = (int)ceil(ri), j = (int)ceil(rj);
    const double ai = ri - (double)(i - 1), aj = rj - (double)(j - 1);
    int ks = 1;
    for (int jj = j - 1; jj <= j; ++jj)
        for (int ii = i - 1; ii <= i; ++ii) {
            int k = 1;
            for (; k <= nlev; ++k) {
                const double z = hgt[(ii - 1) + (long)nx * (jj - 1) + nij * (k - 1)];
                if (z > -300.0 && z < 10000.0) break;
            }
            if (k > ks) ks = k;
        }
    if (ks > nlev) return IQC_OUT_VLO;
    /* itpl_2d_column at (ri, rj), term order of the reference */
    double zl[512];
    if (nlev > 512) return IQC_OUT_VHI;
    for (int k = 1; k <= nlev; ++k) {
        const double *z = hgt + nij * (k - 1);
        const long c00 = (i - 2) + (long)nx * (j - 2), c10 = c00 + 1, c01 = c00 + nx,
                   c11 = c01 + 1;
        zl[k - 1] = z[c00] * (1.0 - ai) * (1.0 - aj) + z[c10] * ai * (1.0 - aj) +
                    z[c01] * (1.0 - ai) * aj + z[c11] * ai * aj;
    }
    if (rlev > zl[nlev - 1]) return IQC_OUT_VHI;
    if (rlev < zl[ks - 1]) return IQC_OUT_VLO;
    int k = ks + 1;
    for (; k <= nlev; ++k)
        if (zl[k - 1] > rlev) break;
    if (k > nlev) k = nlev;
    const double ak = (rlev - zl[k - 2]) / (zl[k - 1] - zl[k - 2]);
    *rk = (double)(k - 1) + ak;
    if (!(*rk >= 1.0 && *rk <= (double)nlev)) {  /* heights not ascending */
        *rk = ORC_UNDEF;
        return IQC_OUT_VLO;
    }
    return IQC_GOOD;
}

/*
 * itpl_3d as Trans_XtoY_radar calls it (scale/common/common_obs_scale.f90:1298-1325
 * with the level index first): trilinear in (level rk, x ri, y rj), 1-based,
 * over var(ij, lev) with stride sl between levels.  A level coordinate of
 * exactly 1 takes layer (1, 2) with weight 0 on level 2, the same terms the
 * reference sums with its halo level weighted 0.
 */
double orc_itpl_radar(const double *var, long sl, int nx, double ri, double rj, double rk)
{
    int kl = (int)ceil(rk);
    if (kl < 2) kl = 2;
    const int ix = (int)ceil(ri), jy = (int)ceil(rj);
    const double al = rk - (double)(kl - 1), ax = ri - (double)(ix - 1), ay = rj - (double)(jy - 1);
    const long c00 = (ix - 2) + (long)nx * (jy - 2), c10 = c00 + 1, c01 = c00 + nx, c11 = c01 + 1;
    const double *lo = var + sl * (kl - 2), *hi = var + sl * (kl - 1);
    return lo[c00] * (1.0 - al) * (1.0 - ax) * (1.0 - ay) + hi[c00] * al * (1.0 - ax) * (1.0 - ay) +
           lo[c10] * (1.0 - al) * ax * (1.0 - ay) + hi[c10] * al * ax * (1.0 - ay) +
           lo[c01] * (1.0 - al) * (1.0 - ax) * ay + hi[c01] * al * (1.0 - ax) * ay +
           lo[c11] * (1.0 - al) * ax * ay + hi[c11] * al * ax * ay;
}

/*
 * calc_ref_vr (scale/common/common_obs_scale.f90:609-960): reflectivity factor
 * ref (mm^6 m^-3) and radial velocity vr from the mixing ratios, wind,
 * temperature and pressure at the obs, for METHOD_REF_CALC 1, 2 or 3.
 * gam[0..2] = Gamma(4.8), Gamma(4.25), Gamma(4.5) (com_gamma of 4 + b,
 * 4 + d and 4.5).  Returns 0, or -1 for an unknown method.
 */
int orc_calc_ref_vr(double qv, double qc, double qr, double qci, double qs, double qg, double u,
                    double v, double w, double t, double p, double az, double elev, int method,
                    int use_wt, const double *gam, double *ref_out, double *vr_out)
{
    (void)qv;
    (void)qc;
    (void)qci;
    const double pi = ORC_PI, deg2rad = ORC_PI / 180.0;
    double zr = 0.0, zs = 0.0, zg = 0.0, zms = 0.0, zmg = 0.0, ref = 0.0, wt = 0.0;
    double ro = p / (ORC_RD * t);
    if (method == 1) {
        const double nor = 8.0e6, ror = 1000.0, pip = pow(pi, 1.75), cf = 10.0e18 * 72, p0 = 1.0e5;
        const double qt = qr + qs + qg;
        if (qt > 0.0) {
            ref = cf * pow(ro * qt, 1.75);
            ref = ref / (pip * pow(nor, 0.75) * pow(ror, 1.75));
            const double a = pow(p0 / p, F4(0.4));
            wt = 5.40 * a * pow(qt, 0.125);
        }
    } else if (method == 2) {
        double nor = 8.0e6, nos = 3.0e6, nog = 4.0e4, ror = 1000.0, ros = 100.0, rog = 913.0,
               roi = 917.0, roo = 1.0;
        const double ki2 = 0.176, kr2 = 0.930, pip = pow(pi, 1.75), cf = 1.0e18 * 720;
        if (qr > 0.0) {
            zr = cf * pow(ro * qr, 1.75);
            zr = zr / (pip * pow(nor, 0.75) * pow(ror, 1.75));
        }
        if (qs > 0.0) {
            if (t <= F4(273.16)) {
                zs = cf * ki2 * pow(ros, 0.25) * pow(ro * qs, 1.75);
                zs = zs / (pip * kr2 * pow(nos, 0.75) * (roi * roi));
            } else {
                zs = cf * pow(ro * qs, 1.75);
                zs = zs / (pip * pow(nos, 0.75) * pow(roi, 1.75));
            }
        }
        if (qg > 0.0) {
            zg = pow(cf / (pip * pow(nog, 0.75) * pow(rog, 1.75)), F4(0.95));
            zg = zg * pow(ro * qg, F4(1.6625));
        }
        ref = zr + zs + zg;
        if (ref > 0.0) {
            nor = nor * F4(1e-3);
            nos = nos * F4(1e-3);
            nog = nog * F4(1e-3);
            ror = ror * F4(1e-3);
            ros = ros * F4(1e-3);
            rog = rog * F4(1e-3);
            roo = roo * F4(1e-3);
            ro = ro * F4(1e-3);
            const double a = 2115.0, b = 0.8, c = 152.93, d = 0.25, Cd = 0.6;
            const double lr = pow(pi * ror * nor / (ro * qr), 0.25);
            const double ls = pow(pi * ros * nos / (ro * qs), 0.25);
            const double lg = pow(pi * rog * nog / (ro * qg), 0.25);
            const double rofactor = pow(roo / ro, 0.25);
            double wr = a * gam[0] / (6.0 * pow(lr, b));
            wr = 1.0e-2 * wr * rofactor;
            double ws = c * gam[1] / (6.0 * pow(ls, d));
            ws = 1.0e-2 * ws * rofactor;
            double wg = gam[2] * pow((4.0 * ORC_GG * 100.0 * rog) / (3.0 * Cd * ro), 0.5);
            wg = 1.0e-2 * wg / (6.0 * pow(lg, 0.5));
            wt = (wr * zr + ws * zs + wg * zg) / (zr + zs + zg);
        }
    } else if (method == 3) {
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
        if (qrp > 0.0) zr = 2.53e4 * pow(ro * qrp * 1.0e3, F4(1.84));
        if (qsp > 0.0) zs = 3.48e3 * pow(ro * qsp * 1.0e3, F4(1.66));
        if (qgp > 0.0) zg = 8.18e4 * pow(ro * qgp * 1.0e3, 1.50);
        if (qms > 0.0) {
            zms = (F4(0.00491) + 5.75 * fws - F4(5.588) * (fws * fws)) * 1.0e5;
            zms = zms * pow(ro * qms * 1.0e3, F4(1.67) - F4(0.202) * fws + F4(0.398) * (fws * fws));
        }
        if (qmg > 0.0) {
            zmg = (F4(0.809) + F4(10.13) * fwg - F4(5.98) * (fwg * fwg)) * 1.0e5;
            zmg = zmg * pow(ro * qmg * 1.0e3, F4(1.48) + F4(0.0448) * fwg - F4(0.0313) * (fwg * fwg));
        }
        ref = zr + zg + zs + zms + zmg;
        if (ref > 0.0) {
            const double nor = 8.0e-2, nos = 3.0e-2, nog = 4.0e-4, ror = 1.0, ros = 0.1, rog = 0.917,
                         roo = 0.001;
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
                wg = gam[2] * pow((4.0 * ORC_GG * 100.0 * rog) / (3.0 * Cd * ro), 0.5);
                wg = 1.0e-2 * wg / (6.0 * pow(lg, 0.5));
            }
            wt = (wr * zr + ws * zs + ws * zms + wg * zg + wg * zmg) / (zr + zs + zg + zms + zmg);
        }
    } else {
        return -1;
    }
    double vr = u * cos(elev * deg2rad) * sin(az * deg2rad);
    vr = vr + v * cos(elev * deg2rad) * cos(az * deg2rad);
    if (use_wt)
        vr = vr + (w - wt) * sin(elev * deg2rad);
    else
        vr = vr + w * sin(elev * deg2rad);
    *ref_out = ref;
    *vr_out = vr;
    return 0;
}

/*
 * The radar operator of SCALE's obsope for one (obs, member): the height and
 * range checks of obsope_tools.f90:353-361 (RADAR_ZMAX, phys2ijkz), then
 * Trans_XtoY_radar (common_obs_scale.f90:338-476, stggrd = 0) and obsope's
 * reset of iqc_ref_low to good (obsope_tools.f90:380).
 * x3: the member's 3-D fields var(ij, lev) for u, v, w, t, p, q, qc, qr, qi,
 * qs, qg at x3 + f * sv (sv = the stride between fields, nij*nlev for a
 * contiguous member); hgt(ij, lev) the level heights.  opt = {min_ref_dbz,
 * low_ref_shift, zmax}.  Writes y and rk; returns the qc.
 */
int orc_radar_xtoy(int elm, double lon, double lat, double lev, double ri, double rj,
                   double radar_lon, double radar_lat, double radar_z, const double *x3, long sv,
                   const double *hgt, int nx, int ny, int nlev, int method, int use_wt,
                   const double *opt, const double *gam, double *y, double *rk)
{
    const double deg2rad = ORC_PI / 180.0, rad2deg = 180.0 / ORC_PI;
    const long sl = (long)nx * ny;
    *y = ORC_UNDEF;
    *rk = ORC_UNDEF;
    if (elm != ID_RADAR_REF && elm != ID_RADAR_VR) return IQC_OTYPE;
    if (lev > opt[2]) return IQC_RADAR_VHI;
    int qc = orc_phys2ijkz(hgt, nx, ny, nlev, ri, rj, lev, rk);
    if (qc != IQC_GOOD) return qc;
    double f[11];
    for (int n = 0; n < 11; ++n) f[n] = orc_itpl_radar(x3 + sv * n, sl, nx, ri, rj, *rk);
    const double dlon = lon - radar_lon, dlat = lat - radar_lat;
    if (dlon == 0.0 && dlat == 0.0) return IQC_OUT_H;
    double az = rad2deg * atan2(dlon * cos(radar_lat * deg2rad), dlat);
    if (az < 0.0) az = 360.0 + az;
    const double dist = orc_distll_1(lon, lat, radar_lon, radar_lat);
    const double elev = rad2deg * atan2(lev - radar_z, dist);
    double ref = 0.0, vr = 0.0;
    if (orc_calc_ref_vr(f[5], f[6], f[7], f[8], f[9], f[10], f[0], f[1], f[2], f[3], f[4], az, elev,
                        method, use_wt, gam, &ref, &vr) != 0)
        return IQC_OTYPE;
    const double min_ref = pow(10.0, opt[0] / 10.0);
    if (elm == ID_RADAR_REF)
        *y = ref < min_ref ? opt[0] + opt[1] : 10.0 * log10(ref);
    else
        *y = vr;
    return IQC_GOOD;  /* iqc_ref_low is reset to good by obsope */
}

/* scale/common/common_obs_scale.f90:63-78,141-151 */
#define ID_SCALE_RAIN 19999
#define ID_RADAR_PRH 4003
#define ID_H08IR 8800
#define ID_TCLON 99991
#define ID_TCLAT 99992
#define ID_TCMIP 99993
#define IQC_GROSS_ERR 5
#define IQC_REF_MEM 12
#define IQC_OBS_BAD 50

/* cfg indices of orc_obs_departure_scale (the namelist values of
 * scale/common/common_nml.f90:83-91,140-146,165-174) */
enum {
    RQ_USE_REF, RQ_USE_VR, RQ_USE_OBSERR_REF, RQ_USE_OBSERR_VR, RQ_MIN_MEMBER,
    RQ_MIN_MEMBER_OBSREF, RQ_THRES_DBZ, RQ_MIN_REF_DBZ, RQ_LOW_REF_SHIFT, RQ_OBSERR_REF,
    RQ_OBSERR_VR, RQ_GROSS, RQ_GROSS_RAIN, RQ_GROSS_REF, RQ_GROSS_VR, RQ_GROSS_PRH,
    RQ_GROSS_TCX, RQ_GROSS_TCY, RQ_GROSS_TCP, RQ_N
};

/*
 * SCALE's departure and QC step of set_letkf_obs for one obs
 * (scale/letkf/letkf_obs.f90):
 *  - preprocessing (:320-343): a reflectivity obs in mm^6 m^-3 in [0, 1e10)
 *    becomes 10 log10(dat) dBZ, or MIN_RADAR_REF_DBZ + LOW_REF_SHIFT below
 *    MIN_RADAR_REF = 10^(MIN_RADAR_REF_DBZ/10) (common_obs_scale.f90:249);
 *    outside that range undef.  USE_OBSERR_RADAR_REF / _VR replace the error;
 *  - the qc of the obs is the maximum of the members' operator qc (the
 *    MPI_MAX all-reduce of :266-272); a positive qc ends the step (:357);
 *  - radar checks (:476-508): USE_RADAR_REF / USE_RADAR_VR false -> otype;
 *    an undef reflectivity -> obs_bad; mem_ref = members with H(x) above
 *    RADAR_REF_THRES_DBZ + 1e-6; an obs above that value needs
 *    MIN_RADAR_REF_MEMBER_OBSREF such members, any other
 *    MIN_RADAR_REF_MEMBER, else ref_mem.  RADAR_REF_THRES_DBZ below
 *    MIN_RADAR_REF_DBZ is raised to it (common_nml.f90:596-598);
 *  - mean of H(x) summed in member order, perturbations, departure
 *    y - mean (:560-576);
 *  - the gross-error check of each element (:590-645); a factor below zero
 *    stands for GROSS_ERROR (common_nml.f90:348-371).
 * qc_mem (ne) the operator's qc of each member (NULL: all good); hx (ne) in
 * H(x), out the perturbations when the returned qc is good (else unchanged).
 * Writes dat, err (preprocessed) and dep (0 unless good).  Returns the qc, or
 * -1 for an H08 obs (its cloud flags come from the RTTOV operator, not built).
 */
int orc_obs_departure_scale(int elm, double dat_in, double err_in, const int *qc_mem, int ne,
                            double *hx, const double *cfg, double *dat, double *err, double *dep)
{
    double thres = cfg[RQ_THRES_DBZ];
    if (thres < cfg[RQ_MIN_REF_DBZ]) thres = cfg[RQ_MIN_REF_DBZ];
    const double min_ref = pow(10.0, cfg[RQ_MIN_REF_DBZ] / 10.0);
    double y = dat_in, e = err_in;
    *dep = 0.0;
    if (elm == ID_H08IR) return -1;
    if (elm == ID_RADAR_REF) {
        if (y >= 0.0 && y < 1.0e10)
            y = y < min_ref ? cfg[RQ_MIN_REF_DBZ] + cfg[RQ_LOW_REF_SHIFT] : 10.0 * log10(y);
        else
            y = ORC_UNDEF;
        if (cfg[RQ_USE_OBSERR_REF] != 0.0) e = cfg[RQ_OBSERR_REF];
    }
    if (elm == ID_RADAR_VR && cfg[RQ_USE_OBSERR_VR] != 0.0) e = cfg[RQ_OBSERR_VR];
    *dat = y;
    *err = e;
    int qc = IQC_GOOD;
    if (qc_mem)
        for (int m = 0; m < ne; ++m)
            if (qc_mem[m] > qc) qc = qc_mem[m];
    if (qc > 0) return qc;
    if (elm == ID_RADAR_REF) {
        if (cfg[RQ_USE_REF] == 0.0) return IQC_OTYPE;
        if (y == ORC_UNDEF) return IQC_OBS_BAD;
        int mem_ref = 0;
        for (int m = 0; m < ne; ++m)
            if (hx[m] > thres + 1.0e-6) ++mem_ref;
        const double need = y > thres + 1.0e-6 ? cfg[RQ_MIN_MEMBER_OBSREF] : cfg[RQ_MIN_MEMBER];
        if (mem_ref < (int)need) return IQC_REF_MEM;
    }
    if (elm == ID_RADAR_VR && cfg[RQ_USE_VR] == 0.0) return IQC_OTYPE;
    double mean = hx[0];
    for (int m = 1; m < ne; ++m) mean = mean + hx[m];
    mean = mean / (double)ne;
    for (int m = 0; m < ne; ++m) hx[m] = hx[m] - mean;
    const double d = y - mean;
    *dep = d;
    int g;
    switch (elm) {
    case ID_SCALE_RAIN: g = RQ_GROSS_RAIN; break;
    case ID_RADAR_REF: g = RQ_GROSS_REF; break;
    case ID_RADAR_VR: g = RQ_GROSS_VR; break;
    case ID_RADAR_PRH: g = RQ_GROSS_PRH; break;
    case ID_TCLON: g = RQ_GROSS_TCX; break;
    case ID_TCLAT: g = RQ_GROSS_TCY; break;
    case ID_TCMIP: g = RQ_GROSS_TCP; break;
    default: g = RQ_GROSS; break;
    }
    const double fac = cfg[g] < 0.0 ? cfg[RQ_GROSS] : cfg[g];
    if (fabs(d) > fac * e) return IQC_GROSS_ERR;
    return IQC_GOOD;
}
