// lekf_solve.hip -- batched LETKF local solve on gfx950.
//
// One 64-lane wavefront per local analysis ("point"):
//   1. gather the point's local obs rows of Y from the obs table and form the
//      Gram matrix G = [w.Y  w.d]^T [Y d] on MFMA (v_mfma_f64_16x16x4f64, or
//      FP32 / split-BF16x3 MFMA in the mixed form).  G holds C = Y^T R^-1 Y,
//      b = Y^T R^-1 d and sum w d^2 in one product
//      (common/common_letkf.f90:101-118, 177-182, 219-227);
//   2. C += (k-1)/alpha I (common_letkf.f90:130-138);
//   3. symmetric eigendecomposition by parallel cyclic Jacobi in LDS
//      (stands in for mtx_eigen / EISPACK rs, common/common_mtx.f90:41-99);
//   4. W assembly: w-bar = V L^-1 V^T b, Wa = V sqrt((k-1)/L) V^T,
//      T = Wa + w-bar 1^T, Pa = V L^-1 V^T (common_letkf.f90:146-214);
//   5. adaptive inflation update (common_letkf.f90:218-239);
//   6. optionally the fused apply Xa = xmean + Xb' T over the point's state
//      rows (speedy/letkf/letkf_tools.f90:94-112,164-177,192-197).
// A point's arithmetic order depends only on its own data, so its result is
// the same whatever else shares the launch.
#include "lekf_dev.hpp"

namespace {

// Circle-method pairing of n (even) indices: round r in [0, n-1), pair i in
// [0, n/2).  Over the n-1 rounds every index pair meets exactly once.
__device__ __forceinline__ void rr_pair(int n, int r, int i, int& p, int& q)
{
    const int m = n - 1;
    int a, b;
    if (i == 0) {
        a = 0;
        b = 1 + (r % m);
    } else {
        a = 1 + (i + r) % m;
        b = 1 + (m - i + r) % m;
    }
    p = a < b ? a : b;
    q = a < b ? b : a;
}

struct SolveLds {
    double* A;    // n x lda   (matrix being diagonalised, later T)
    double* V;    // k x lda   (eigenvectors, columns)
    double* C;    // np rotation cosines
    double* S;    // np sines
    double* T;    // np tangents
    int* P;       // np
    int* Q;       // np
    int* blk;     // nblk packed (i << 16 | j)
    double* B;    // k   (Y^T R^-1 d)
    double* L;    // n   eigenvalues
    double* U;    // k
    double* Wb;   // k   w-bar
    double* X;    // LEKF_NVMAX * k   state perturbation rows (aliases V)
    double* M;    // LEKF_NVMAX       row means (aliases V)
    double* sc;   // 8 scalars
    int* flag;    // 2 ints
};

__host__ __device__ inline size_t solve_lds_doubles(int k)
{
    const int n = (k + 1) & ~1, lda = n + 1, np = n / 2;
    const int nblk = np * (np + 1) / 2;
    const size_t vx = (size_t)LEKF_NVMAX * k + LEKF_NVMAX;  // apply rows reuse the V region
    size_t d = 0;
    d += (size_t)n * lda;                                       // A
    d += (size_t)k * lda > vx ? (size_t)k * lda : vx;           // V, later X and M
    d += 3 * (size_t)np;                                        // C S T
    d += (size_t)np;                                            // P,Q (2 ints = 1 double each pair)
    d += ((size_t)nblk + 1) / 2;                                // blk ints
    d += (size_t)k + n + 2 * (size_t)k;                         // B L U Wb
    d += 8 + 1;                                                 // scalars, flag
    return d;
}

__device__ inline SolveLds carve(double* s, int k)
{
    const int n = (k + 1) & ~1, lda = n + 1, np = n / 2;
    const int nblk = np * (np + 1) / 2;
    SolveLds L;
    const size_t vx = (size_t)LEKF_NVMAX * k + LEKF_NVMAX;
    L.A = s;
    s += (size_t)n * lda;
    L.V = s;
    L.X = s;                      // used only after T has been formed
    L.M = s + (size_t)LEKF_NVMAX * k;
    s += (size_t)k * lda > vx ? (size_t)k * lda : vx;
    L.C = s;
    s += np;
    L.S = s;
    s += np;
    L.T = s;
    s += np;
    L.P = (int*)s;
    L.Q = L.P + np;
    s += np;  // np doubles hold the 2*np ints of P and Q
    L.blk = (int*)s;
    s += (nblk + 1) / 2;
    L.B = s;
    s += k;
    L.L = s;
    s += n;
    L.U = s;
    s += k;
    L.Wb = s;
    s += k;
    L.sc = s;
    s += 8;
    L.flag = (int*)s;
    return L;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Reciprocal and reciprocal square root: hardware estimate refined by Newton
// steps to full FP64 accuracy (a Jacobi rotation only needs c^2 + s^2 = 1 to
// rounding, which an accurate c = (1+t^2)^-1/2 and s = t c give).
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

__device__ __forceinline__ double rsq_nr(double x)
{
    double r = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    r = r * fma(-h * r, r, 1.5);
    r = r * fma(-h * r, r, 1.5);
    return r;
}

// ---------------------------------------------------------------------------
// Gram products.  Lane l holds A[i = l%16][kk = l/16] and B[kk][j = l%16] of
// each 16x16x4 step (16x16x32 BF16: kk = 8*(l/16) .. +7).
// ---------------------------------------------------------------------------
template <int NT>
__device__ __forceinline__ void gram_f64(const LekfJob& J, long off, int p, int k, int lane,
                                         lekf_d4* acc, double& p3, int wv, int nw)
{
    const int lr = lane & 15, kq = lane >> 4;
    const bool wloc = (J.flags & LEKF_F_WLOC) != 0;
#pragma unroll 2
    for (int q0 = 4 * wv; q0 < p; q0 += 4 * nw) {
        const int q = q0 + kq;
        const bool valid = q < p;
        double w = 0.0, dd = 0.0;
        long o = 0;
        if (valid) {
            o = J.loc_idx[off + q];
            const double rho = J.loc_rho[off + q];
            const double rd = J.loc_rdiag ? J.loc_rdiag[off + q] : J.rdiag[o];
            w = (wloc ? 1.0 : rho) / rd;
            dd = J.dep[o];
            if (lr == 0) p3 += rho;
        }
        double v[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int c = 16 * t + lr;
            double x = 0.0;
            if (valid) {
                if (c < k)
                    x = J.hdxf[o * J.y_so + (long)c * J.y_sm];
                else if (c == k)
                    x = dd;
            }
            v[t] = x;
        }
        int a = 0;
#pragma unroll
        for (int I = 0; I < NT; ++I) {
            const double ai = w * v[I];
#pragma unroll
            for (int Jt = I; Jt < NT; ++Jt) {
                acc[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, v[Jt], acc[a], 0, 0, 0);
                ++a;
            }
        }
    }
}

// Mixed form: Y from the FP32 table, products accumulated in FP32 on MFMA.
// FP32: v_mfma_f32_16x16x4f32 (operands rounded to FP32).
// BF16x3: each FP32 operand split x = h + l (h = bf16(x), l = bf16(x - h));
// A*B ~ Ah*Bh + Ah*Bl + Al*Bh on v_mfma_f32_16x16x32_bf16.
template <int NT>
__device__ __forceinline__ void gram_f32(const LekfJob& J, long off, int p, int k, int lane,
                                         lekf_f4* acc, double& p3, int wv, int nw)
{
    const int lr = lane & 15, kq = lane >> 4;
    const bool wloc = (J.flags & LEKF_F_WLOC) != 0;
#pragma unroll 2
    for (int q0 = 4 * wv; q0 < p; q0 += 4 * nw) {
        const int q = q0 + kq;
        const bool valid = q < p;
        float w = 0.0f, dd = 0.0f;
        long o = 0;
        if (valid) {
            o = J.loc_idx[off + q];
            const double rho = J.loc_rho[off + q];
            const double rd = J.loc_rdiag ? J.loc_rdiag[off + q] : J.rdiag[o];
            w = (float)((wloc ? 1.0 : rho) / rd);
            dd = (float)J.dep[o];
            if (lr == 0) p3 += rho;
        }
        float v[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int c = 16 * t + lr;
            float x = 0.0f;
            if (valid) {
                if (c < k)
                    x = J.hdxf32[o * (long)k + c];
                else if (c == k)
                    x = dd;
            }
            v[t] = x;
        }
        int a = 0;
#pragma unroll
        for (int I = 0; I < NT; ++I) {
            const float ai = w * v[I];
#pragma unroll
            for (int Jt = I; Jt < NT; ++Jt) {
                acc[a] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, v[Jt], acc[a], 0, 0, 0);
                ++a;
            }
        }
    }
}

// Split-BF16x3: K = 32 obs per MFMA step.  Lane l supplies 8 consecutive k
// values (kk = 8*(l/16) .. +7) of row/col l%16.
template <int NT>
__device__ __forceinline__ void gram_bf16x3(const LekfJob& J, long off, int p, int k, int lane,
                                            lekf_f4* acc, double& p3, int wv, int nw)
{
    const int lr = lane & 15, kq = lane >> 4;
    const bool wloc = (J.flags & LEKF_F_WLOC) != 0;
    for (int q0 = 32 * wv; q0 < p; q0 += 32 * nw) {
        lekf_bf8 ah[NT], al[NT], bh[NT], bl[NT];
        float wv[8];
        long ov[8];
        bool vv[8];
        float dv[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int q = q0 + 8 * kq + e;
            vv[e] = q < p;
            wv[e] = 0.0f;
            dv[e] = 0.0f;
            ov[e] = 0;
            if (vv[e]) {
                const long o = J.loc_idx[off + q];
                const double rho = J.loc_rho[off + q];
                const double rd = J.loc_rdiag ? J.loc_rdiag[off + q] : J.rdiag[o];
                wv[e] = (float)((wloc ? 1.0 : rho) / rd);
                dv[e] = (float)J.dep[o];
                ov[e] = o;
                if (lr == 0) p3 += rho;
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int c = 16 * t + lr;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                float x = 0.0f;
                if (vv[e]) {
                    if (c < k)
                        x = J.hdxf32[ov[e] * (long)k + c];
                    else if (c == k)
                        x = dv[e];
                }
                const float xa = wv[e] * x;
                const __bf16 bh1 = (__bf16)x;
                const __bf16 bl1 = (__bf16)(x - (float)bh1);
                const __bf16 ah1 = (__bf16)xa;
                const __bf16 al1 = (__bf16)(xa - (float)ah1);
                bh[t][e] = bh1;
                bl[t][e] = bl1;
                ah[t][e] = ah1;
                al[t][e] = al1;
            }
        }
        int a = 0;
#pragma unroll
        for (int I = 0; I < NT; ++I) {
#pragma unroll
            for (int Jt = I; Jt < NT; ++Jt) {
                acc[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[I], bh[Jt], acc[a], 0, 0, 0);
                acc[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[I], bl[Jt], acc[a], 0, 0, 0);
                acc[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[I], bh[Jt], acc[a], 0, 0, 0);
                ++a;
            }
        }
    }
}

// Write the upper tiles of G into LDS: C (k x k, exactly symmetric), b and
// sum w d^2; with add, accumulate onto what an earlier wave stored.  Only
// (row <= col) elements are taken so C is symmetric bit for bit.  Result-tile
// layouts (measured on gfx950): the FP64 16x16x4 MFMA puts element
// (row = l/16 + 4r, col = l%16) in register r of lane l; the FP32 and BF16
// 16x16 forms put (row = 4*(l/16) + r, col = l%16) there.
template <int NT, bool F64, typename ACC>
__device__ __forceinline__ void store_gram(const ACC* acc, int k, int lda, int lane, SolveLds& L,
                                           bool add)
{
    const int lr = lane & 15, kq = lane >> 4;
    int a = 0;
#pragma unroll
    for (int I = 0; I < NT; ++I) {
#pragma unroll
        for (int Jt = I; Jt < NT; ++Jt) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * I + (F64 ? kq + 4 * r : 4 * kq + r);
                const int col = 16 * Jt + lr;
                const double val = (double)acc[a][r];
                if (row <= col) {
                    if (col < k) {
                        const double v = add ? L.A[row * lda + col] + val : val;
                        L.A[row * lda + col] = v;
                        L.A[col * lda + row] = v;
                    } else if (col == k) {
                        if (row < k)
                            L.B[row] = add ? L.B[row] + val : val;
                        else
                            L.sc[0] = add ? L.sc[0] + val : val;  // sum w d^2
                    }
                }
            }
            ++a;
        }
    }
}

// Parallel cyclic Jacobi on the n x n symmetric matrix in L.A (dummy index
// k when k is odd stays zero and never rotates) by a block of BS threads.
// V accumulates the rotations.  Returns the number of sweeps, or -1 if the
// limit was reached.
template <int BS>
__device__ int jacobi_lds(SolveLds& L, int k, int lane)
{
    const int n = (k + 1) & ~1, lda = n + 1, np = n / 2;
    const int nblk = np * (np + 1) / 2;
    const int dr = BS / np, dj = BS % np;
    int sweep = 0;
    for (; sweep < LEKF_JACOBI_MAX_SWEEPS; ++sweep) {
        if (lane == 0) L.flag[0] = 0;
        __syncthreads();
        for (int r = 0; r < n - 1; ++r) {
            if (lane < np) {
                int pp, qq;
                rr_pair(n, r, lane, pp, qq);
                double c = 1.0, s = 0.0, t = 0.0;
                const double app = L.A[pp * lda + pp];
                const double aqq = L.A[qq * lda + qq];
                const double apq = L.A[pp * lda + qq];
                if (apq * apq > (LEKF_JACOBI_TOL * LEKF_JACOBI_TOL) * fabs(app * aqq)) {
                    const double th = (aqq - app) * rcp_nr(2.0 * apq);
                    const double q = fma(th, th, 1.0);
                    const double at = fabs(th) + q * rsq_nr(q);
                    t = th >= 0.0 ? rcp_nr(at) : -rcp_nr(at);
                    c = rsq_nr(fma(t, t, 1.0));
                    s = t * c;
                    L.flag[0] = 1;
                }
                L.C[lane] = c;
                L.S[lane] = s;
                L.T[lane] = t;
                L.P[lane] = pp;
                L.Q[lane] = qq;
            }
            __syncthreads();
            // A <- R^T A R on 2x2 blocks of rotation pairs (i <= j), mirrored
            for (int w = lane; w < nblk; w += BS) {
                const int e = L.blk[w];
                const int i = e >> 16, j = e & 0xffff;
                const double si = L.S[i];
                const double sj = L.S[j];
                if (si == 0.0 && sj == 0.0) continue;
                const int pi = L.P[i], qi = L.Q[i];
                if (i == j) {
                    const double t = L.T[i];
                    const double apq = L.A[pi * lda + qi];
                    L.A[pi * lda + pi] -= t * apq;
                    L.A[qi * lda + qi] += t * apq;
                    L.A[pi * lda + qi] = 0.0;
                    L.A[qi * lda + pi] = 0.0;
                } else {
                    const double ci = L.C[i], cj = L.C[j];
                    const int pj = L.P[j], qj = L.Q[j];
                    const double b00 = L.A[pi * lda + pj], b01 = L.A[pi * lda + qj];
                    const double b10 = L.A[qi * lda + pj], b11 = L.A[qi * lda + qj];
                    const double x0 = cj * b00 - sj * b01, x1 = sj * b00 + cj * b01;
                    const double y0 = cj * b10 - sj * b11, y1 = sj * b10 + cj * b11;
                    const double n00 = ci * x0 - si * y0, n10 = si * x0 + ci * y0;
                    const double n01 = ci * x1 - si * y1, n11 = si * x1 + ci * y1;
                    L.A[pi * lda + pj] = n00;
                    L.A[pi * lda + qj] = n01;
                    L.A[qi * lda + pj] = n10;
                    L.A[qi * lda + qj] = n11;
                    L.A[pj * lda + pi] = n00;
                    L.A[qj * lda + pi] = n01;
                    L.A[pj * lda + qi] = n10;
                    L.A[qj * lda + qi] = n11;
                }
            }
            // V <- V R (rows r < k, pair j)
            {
                int rr = lane / np, j = lane % np;
                for (; rr < k;) {
                    const double s = L.S[j];
                    if (s != 0.0) {
                        const double c = L.C[j];
                        const int pj = L.P[j], qj = L.Q[j];
                        const double v0 = L.V[rr * lda + pj], v1 = L.V[rr * lda + qj];
                        L.V[rr * lda + pj] = c * v0 - s * v1;
                        L.V[rr * lda + qj] = s * v0 + c * v1;
                    }
                    rr += dr;
                    j += dj;
                    if (j >= np) {
                        j -= np;
                        ++rr;
                    }
                }
            }
            __syncthreads();
        }
        if (L.flag[0] == 0) break;
        __syncthreads();
    }
    return sweep < LEKF_JACOBI_MAX_SWEEPS ? sweep : -1;
}

// Partner of index j in round r of the circle-method tournament on n (even)
// indices: the same pairing as rr_pair, inverted.
__device__ __forceinline__ int rr_partner(int n, int r, int j)
{
    const int m = n - 1;
    if (j == 0) return 1 + (r % m);
    const int x = (j - 1 - (r % m) + m) % m;
    return x == 0 ? 0 : 1 + ((m - x) + r) % m;
}

// One-sided (Hestenes) cyclic Jacobi with the matrix in registers, for k <= 32.
// Lanes 0..31 hold the columns of A (initially C, so A V -> U S with S = the
// eigenvalues of the SPD C); lanes 32..63 hold the columns of V (initially I).
// Each round pairs every column with its tournament partner; a lane fetches
// its partner's column with a cross-lane shuffle, the A lanes form
// [alpha gamma; gamma beta] = the 2x2 Gram of the pair and the rotation that
// zeroes gamma (the same rotation two-sided Jacobi would apply to C'C), and
// both halves apply it to their own column.  KR = rows carried (k rounded up
// to even), rows and columns >= k are zero and never rotate.
// Writes the eigenvalues to L.L and V to L.V (k x k, lda).  Returns the
// number of sweeps or -1 if the sweep limit was reached.
template <int KR>
__device__ int jacobi_reg32(SolveLds& L, int k, int lane, int lda)
{
    const int half = lane >> 5, j = lane & 31;
    double x[KR];
#pragma unroll
    for (int i = 0; i < KR; ++i) {
        if (half == 0)
            x[i] = (j < k && i < k) ? L.A[i * lda + j] : 0.0;
        else
            x[i] = (i == j && j < k) ? 1.0 : 0.0;
    }
    double alpha = 0.0;
#pragma unroll
    for (int i = 0; i < KR; ++i) alpha += x[i] * x[i];
    int sweep = 0;
    for (; sweep < LEKF_JACOBI_MAX_SWEEPS; ++sweep) {
        bool rotated = false;
        for (int r = 0; r < KR - 1; ++r) {
            const int pj = rr_partner(KR, r, j);
            const int src = (half << 5) | pj;
            double y[KR];
#pragma unroll
            for (int i = 0; i < KR; ++i) y[i] = __shfl(x[i], src, 64);
            double gamma = 0.0;
#pragma unroll
            for (int i = 0; i < KR; ++i) gamma += x[i] * y[i];
            const double beta = __shfl(alpha, pj, 64);
            const bool lo = j < pj;
            const double app = lo ? alpha : beta, aqq = lo ? beta : alpha;
            double c = 1.0, s = 0.0, t = 0.0;
            const bool rot =
                gamma * gamma > (LEKF_JACOBI_TOL * LEKF_JACOBI_TOL) * (app * aqq);
            if (rot) {
                const double th = (aqq - app) * rcp_nr(2.0 * gamma);
                const double q = fma(th, th, 1.0);
                const double at = fabs(th) + q * rsq_nr(q);
                t = th >= 0.0 ? rcp_nr(at) : -rcp_nr(at);
                c = rsq_nr(fma(t, t, 1.0));
                s = t * c;
            }
            // the V lanes take the rotation of their column from the A lane
            c = __shfl(c, j, 64);
            s = __shfl(s, j, 64);
            const double sg = lo ? -s : s;
#pragma unroll
            for (int i = 0; i < KR; ++i) x[i] = c * x[i] + sg * y[i];
            if (rot) {
                alpha = lo ? app - t * gamma : aqq + t * gamma;
                rotated = true;
            }
        }
        if (__ballot(rotated && half == 0) == 0ull) break;
        // refresh the column norms once per sweep
        alpha = 0.0;
#pragma unroll
        for (int i = 0; i < KR; ++i) alpha += x[i] * x[i];
    }
    double nrm = 0.0;
#pragma unroll
    for (int i = 0; i < KR; ++i) nrm += x[i] * x[i];
    if (half == 0 && j < k) L.L[j] = sqrt(nrm);
    if (half == 1 && j < k) {
#pragma unroll
        for (int i = 0; i < KR; ++i)
            if (i < k) L.V[i * lda + j] = x[i];
    }
    return sweep < LEKF_JACOBI_MAX_SWEEPS ? sweep : -1;
}

// NT: 16-wide Gram tiles; KR: rows of the register Jacobi (k <= KR <= 32), or
// 0 for the LDS Jacobi (k > 32); MODE 0: FP64 Gram, 1: FP32, 2: split-BF16x3.
// The register-Jacobi classes run one wave per point; the LDS-Jacobi classes
// run four: the Gram is split over the waves by obs and summed into LDS in
// wave order, and the Jacobi, W assembly and apply stride over all threads.
template <int NT, int KR, int MODE>
__global__ __launch_bounds__(KR > 0 ? 64 : 256) void lekf_solve_kernel(
    const LekfJob* __restrict__ jobs, int njobs)
{
    constexpr int BS = KR > 0 ? 64 : 256;  // threads per point
    constexpr int NW = BS / LEKF_WAVE;
    extern __shared__ double smem[];
    const long gp = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & (LEKF_WAVE - 1), wv = tid / LEKF_WAVE;
    const int ji = lekf_job_of(jobs, njobs, gp);
    const LekfJob& J = jobs[ji];
    const int b = (int)(gp - J.pt_begin + J.pt0);
    const int k = J.k;
    const int n = (k + 1) & ~1, lda = n + 1, np = n / 2;
    const int nblk = np * (np + 1) / 2;
    SolveLds L = carve(smem, k);

    const long off = J.loc_off[b];
    const int p = J.loc_cnt ? J.loc_cnt[b] : (int)(J.loc_off[b + 1] - off);
    double alpha = J.infl_in ? J.infl_in[b] : J.infl_fixed;
    const bool transm_mode = (J.flags & LEKF_F_TRANSM) != 0;
    int status = 0;

    // LDS Jacobi (k > 32): zero A (dummy row/col), V = I, block table
    constexpr bool reg_jacobi = KR > 0;
    if (!reg_jacobi) {
        for (int w = tid; w < n * lda; w += BS) L.A[w] = 0.0;
        for (int w = tid; w < k * lda; w += BS) {
            const int r = w / lda, c = w - r * lda;
            L.V[w] = (r == c) ? 1.0 : 0.0;
        }
        for (int w = tid; w < nblk; w += BS) {
            int i = 0, rem = w;
            while (rem >= np - i) {
                rem -= np - i;
                ++i;
            }
            L.blk[w] = (i << 16) | (i + rem);
        }
        __syncthreads();
    }

    if (p > 0) {
        double p3 = 0.0;
        if (MODE == 0) {
            lekf_d4 acc[NT * (NT + 1) / 2];
#pragma unroll
            for (int a = 0; a < NT * (NT + 1) / 2; ++a) acc[a] = lekf_d4{0.0, 0.0, 0.0, 0.0};
            gram_f64<NT>(J, off, p, k, lane, acc, p3, wv, NW);
            for (int w = 0; w < NW; ++w) {
                if (wv == w) store_gram<NT, true>(acc, k, lda, lane, L, w > 0);
                if (NW > 1) __syncthreads();
            }
        } else {
            lekf_f4 acc[NT * (NT + 1) / 2];
#pragma unroll
            for (int a = 0; a < NT * (NT + 1) / 2; ++a) acc[a] = lekf_f4{0.f, 0.f, 0.f, 0.f};
            if (MODE == 1)
                gram_f32<NT>(J, off, p, k, lane, acc, p3, wv, NW);
            else
                gram_bf16x3<NT>(J, off, p, k, lane, acc, p3, wv, NW);
            for (int w = 0; w < NW; ++w) {
                if (wv == w) store_gram<NT, false>(acc, k, lda, lane, L, w > 0);
                if (NW > 1) __syncthreads();
            }
        }
        p3 = wave_sum(p3);
        if (lane == 0) L.sc[4 + wv] = p3;
        __syncthreads();
        if (tid == 0) {
            p3 = L.sc[4];
            for (int w = 1; w < NW; ++w) p3 += L.sc[4 + w];
            double tr = 0.0;
            for (int i = 0; i < k; ++i) tr += L.A[i * lda + i];
            L.sc[1] = tr / (double)(k - 1);  // P2
            L.sc[2] = p3;                    // P3
            if ((J.flags & LEKF_F_MINFL) && J.minfl > 0.0 && alpha < J.minfl) alpha = J.minfl;
            L.sc[3] = alpha;
            const double rho = 1.0 / alpha;
            for (int i = 0; i < k; ++i) L.A[i * lda + i] += (double)(k - 1) * rho;
        }
        __syncthreads();
        alpha = L.sc[3];
        int sweeps;
        if constexpr (reg_jacobi) {
            sweeps = jacobi_reg32<(KR > 0 ? KR : 4)>(L, k, lane, lda);
        } else {
            sweeps = jacobi_lds<BS>(L, k, tid);
            __syncthreads();
            for (int j = tid; j < k; j += BS) L.L[j] = L.A[j * lda + j];
        }
        if (sweeps < 0) status |= LEKF_PT_NOCONV;
        __syncthreads();
        // eigenvalues; reference truncation test (common_mtx.f90:66-74)
        double lmax = 0.0, lmin = 1.0e308;
        for (int j = 0; j < k; ++j) {
            const double lj = L.L[j];
            lmax = lj > lmax ? lj : lmax;
            lmin = lj < lmin ? lj : lmin;
        }
        if (!(lmax > 0.0) || lmin < fabs(lmax) * 1.4901161193847656e-08) status |= LEKF_PT_RANKDEF;
        // u = L^-1 V^T b ; w-bar = V u
        for (int j = tid; j < k; j += BS) {
            double s = 0.0;
            for (int i = 0; i < k; ++i) s += L.V[i * lda + j] * L.B[i];
            L.U[j] = s / L.L[j];
        }
        __syncthreads();
        for (int i = tid; i < k; i += BS) {
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += L.V[i * lda + j] * L.U[j];
            L.Wb[i] = s;
        }
        // U <- sqrt((k-1)/lambda) (U no longer needed)
        __syncthreads();
        for (int j = tid; j < k; j += BS) L.U[j] = sqrt((double)(k - 1) / L.L[j]);
        __syncthreads();
        // T (into A) = V diag(U) V^T (+ w-bar)
        for (int w = tid; w < k * k; w += BS) {
            const int i = w / k, m = w - i * k;
            double s = 0.0;
            for (int l = 0; l < k; ++l) s += L.V[i * lda + l] * L.U[l] * L.V[m * lda + l];
            if (!transm_mode) s += L.Wb[i];
            L.A[i * lda + m] = s;
        }
        if (J.pao) {
            double* pa = J.pao + (size_t)b * k * k;
            for (int w = tid; w < k * k; w += BS) {
                const int i = w / k, m = w - i * k;
                double s = 0.0;
                for (int l = 0; l < k; ++l) s += L.V[i * lda + l] / L.L[l] * L.V[m * lda + l];
                pa[i + (size_t)m * k] = s;
            }
        }
        if (J.transm)
            for (int i = tid; i < k; i += BS) J.transm[(size_t)b * k + i] = L.Wb[i];
        // adaptive inflation (common_letkf.f90:218-239)
        if (tid == 0 && J.infl_out) {
            double na = alpha;
            if (J.flags & LEKF_F_FIXED_INFL) {
                na = J.infl_fixed;
            } else {
                const double p1 = L.sc[0], p2 = L.sc[1], p3v = L.sc[2];
                const double p4 = (p1 - p3v) / p2 - alpha;
                const double qv = (alpha * p2 + p3v) / p2;
                const double so = 2.0 / p3v * (qv * qv);
                const double sb2 = 0.04 * 0.04;
                na = alpha + sb2 / (so + sb2) * p4;
            }
            J.infl_out[b] = na;
        }
    } else {
        // no local obs (common_letkf.f90:82-96)
        const double sa = sqrt(alpha);
        for (int w = tid; w < k * k; w += BS) {
            const int i = w / k, m = w - i * k;
            L.A[i * lda + m] = (i == m) ? sa : 0.0;
        }
        if (J.pao) {
            double* pa = J.pao + (size_t)b * k * k;
            for (int w = tid; w < k * k; w += BS) {
                const int i = w / k, m = w - i * k;
                pa[i + (size_t)m * k] = (i == m) ? alpha / (double)(k - 1) : 0.0;
            }
        }
        if (J.transm)
            for (int i = tid; i < k; i += BS) J.transm[(size_t)b * k + i] = 0.0;
        if (tid == 0 && J.infl_out)
            J.infl_out[b] = (J.flags & LEKF_F_FIXED_INFL) ? J.infl_fixed : alpha;
    }
    if (tid == 0 && J.status) J.status[b] = status;
    __syncthreads();
    if (J.trans) {
        double* tr = J.trans + (size_t)b * k * k;
        for (int w = tid; w < k * k; w += BS) {
            const int m = w / k, i = w - m * k;  // column-major: coalesced over i
            tr[i + (size_t)m * k] = L.A[i * lda + m];
        }
    }
    if (!(J.flags & LEKF_F_APPLY)) return;
    // fused apply: rows of the point (3-D at every level, 2-D at level 0)
    const int lev = b / J.nij, ij = b - lev * J.nij;
    const int n3 = J.nv3d, n2 = (lev == 0) ? J.nv2d : 0;
    const int nrow = n3 + n2;
    const uint32_t mask = J.upd_mask ? J.upd_mask[lev] : 0xffffffffu;
    for (int w = tid; w < nrow * k; w += BS) {
        const int r = w / k, m = w - r * k;
        const double* src = (r < n3) ? J.x3b + ((size_t)b * n3 + r) * k
                                     : J.x2b + ((size_t)ij * J.nv2d + (r - n3)) * k;
        L.X[r * k + m] = src[m];
    }
    __syncthreads();
    // ensemble mean in member order (speedy/common/common_speedy.f90:335-370)
    for (int r = tid; r < nrow; r += BS) {
        double s = L.X[r * k];
        for (int m = 1; m < k; ++m) s += L.X[r * k + m];
        L.M[r] = s / (double)k;
    }
    __syncthreads();
    for (int w = tid; w < nrow * k; w += BS) L.X[w] -= L.M[w / k];
    __syncthreads();
    for (int w = tid; w < nrow * k; w += BS) {
        const int r = w / k, m = w - r * k;
        const bool upd = (r >= n3) || ((mask >> r) & 1u);
        double a = L.M[r];
        if (upd) {
            for (int j = 0; j < k; ++j) a += L.X[r * k + j] * L.A[j * lda + m];
        } else {
            a += L.X[r * k + m];
        }
        double* dst = (r < n3) ? J.x3a + ((size_t)b * n3 + r) * k
                               : J.x2a + ((size_t)ij * J.nv2d + (r - n3)) * k;
        dst[m] = a;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host-side launcher (called from lekf_capi.hip)
// ---------------------------------------------------------------------------
size_t lekf_solve_lds_bytes(int k) { return solve_lds_doubles(k) * sizeof(double); }

int lekf_solve_nt(int k) { return (k + 1 + 15) / 16; }

// Kernel class of an ensemble size: KR * 8 + NT.  k <= 32 uses the register
// Jacobi with KR = k rounded up to a multiple of 4; larger k the LDS Jacobi.
int lekf_solve_class(int k)
{
    if (k <= 32) {
        const int kr = (k + 3) & ~3;
        const int nt = kr <= 12 ? 1 : (kr <= 28 ? 2 : 3);
        return kr * 8 + nt;
    }
    return lekf_solve_nt(k);
}

template <int NT, int KR, int MODE>
static hipError_t launch_k(const LekfJob* djobs, int njobs, long npt, int kmax, hipStream_t st)
{
    const size_t lds = lekf_solve_lds_bytes(kmax);
    hipLaunchKernelGGL((lekf_solve_kernel<NT, KR, MODE>), dim3((unsigned)npt),
                       dim3(KR > 0 ? LEKF_WAVE : 4 * LEKF_WAVE), lds, st, djobs, njobs);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_mode(int cls, const LekfJob* djobs, int njobs, long npt, int kmax,
                              hipStream_t st)
{
    switch (cls) {
    case 4 * 8 + 1: return launch_k<1, 4, MODE>(djobs, njobs, npt, kmax, st);
    case 8 * 8 + 1: return launch_k<1, 8, MODE>(djobs, njobs, npt, kmax, st);
    case 12 * 8 + 1: return launch_k<1, 12, MODE>(djobs, njobs, npt, kmax, st);
    case 16 * 8 + 2: return launch_k<2, 16, MODE>(djobs, njobs, npt, kmax, st);
    case 20 * 8 + 2: return launch_k<2, 20, MODE>(djobs, njobs, npt, kmax, st);
    case 24 * 8 + 2: return launch_k<2, 24, MODE>(djobs, njobs, npt, kmax, st);
    case 28 * 8 + 2: return launch_k<2, 28, MODE>(djobs, njobs, npt, kmax, st);
    case 32 * 8 + 3: return launch_k<3, 32, MODE>(djobs, njobs, npt, kmax, st);
    case 3: return launch_k<3, 0, MODE>(djobs, njobs, npt, kmax, st);
    case 4: return launch_k<4, 0, MODE>(djobs, njobs, npt, kmax, st);
    case 5: return launch_k<5, 0, MODE>(djobs, njobs, npt, kmax, st);
    case 6: return launch_k<6, 0, MODE>(djobs, njobs, npt, kmax, st);
    default: return hipErrorInvalidValue;
    }
}

// Launch the solve over npt flat points of njobs jobs (device array djobs).
// Every job of a launch has the same kernel class; kmax (the largest k) sets
// the LDS size.
hipError_t lekf_solve_launch(const LekfJob* djobs, int njobs, long npt, int kmax, int mode,
                             hipStream_t st)
{
    if (npt <= 0) return hipSuccess;
    const int cls = lekf_solve_class(kmax);
    if (mode == 0) return launch_mode<0>(cls, djobs, njobs, npt, kmax, st);
    if (mode == 1) return launch_mode<1>(cls, djobs, njobs, npt, kmax, st);
    return launch_mode<2>(cls, djobs, njobs, npt, kmax, st);
}
