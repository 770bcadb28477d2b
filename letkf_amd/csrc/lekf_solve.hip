// lekf_solve.hip -- batched LETKF local solve on gfx950.
//
// A local analysis ("point") is one letkf_core solve plus its apply:
//   1. gather the point's local obs rows of Y from the obs table and form the
//      Gram matrix G = [w.Y  w.d]^T [Y d] on MFMA (v_mfma_f64_16x16x4f64, or
//      FP32 / split-BF16x3 MFMA in the mixed form).  G holds C = Y^T R^-1 Y,
//      b = Y^T R^-1 d and sum w d^2 in one product
//      (common/common_letkf.f90:101-118, 177-182, 219-227);
//   2. C += (k-1)/alpha I (common_letkf.f90:130-138);
//   3. symmetric eigendecomposition in FP64 (stands in for mtx_eigen / EISPACK
//      rs, common/common_mtx.f90:41-99);
//   4. W assembly: w-bar = V L^-1 V^T b, Wa = V sqrt((k-1)/L) V^T,
//      T = Wa + w-bar 1^T, Pa = V L^-1 V^T (common_letkf.f90:146-214);
//   5. adaptive inflation update (common_letkf.f90:218-239);
//   6. optionally the fused apply Xa = xmean + Xb' T over the point's state
//      rows (speedy/letkf/letkf_tools.f90:94-112,164-177,192-197).
//
// Two kernel families, chosen by the ensemble size k:
//   * k <= 32: one fused kernel, one wave per point; the eigendecomposition is
//     a one-sided Jacobi with the matrix in registers.
//   * 32 < k <= LEKF_KSPLIT: a Gram kernel (2-8 waves per point, MFMA) writes
//     C, b and the sums to a workspace; a second kernel (one wave per point)
//     reduces C to tridiagonal form by Householder reflections, diagonalises
//     it by implicit QL with shifts (the method of EISPACK tred2/tql2, the
//     reference's own eigensolver), and assembles W and the apply, with the
//     matrix in registers (k <= 64) or LDS.
//   * LEKF_KSPLIT < k <= LEKF_KMAX: a block Gram kernel (one wave per 64 x 64
//     block of a point) writes the same workspace record, and the eigen
//     kernel runs the same steps with the matrix left in that record in
//     global memory (a k = 256 FP64 matrix is 512 KB, beyond the LDS).
// A point's arithmetic depends only on its own data, so its result is the
// same whatever else shares the launch.
#include "lekf_dev.hpp"
#include "lekf_tune.hpp"
#include "lekf_ktime.hpp"

namespace {

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// value of x in lane src (src uniform across the wave)
__device__ __forceinline__ double lane_bcast(double x, int src)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
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
//
// The local list is read in chunks of 64 entries: lane l loads entry q0 + l
// (obs index, weight rho/rdiag, departure) with one coalesced access, and the
// MFMA steps of the chunk take them from the owning lanes by shuffles.  All
// gathers of the chunk's Y rows are then independent loads, issued together,
// instead of one dependent index -> row chain per step.  Wave wv of nw takes
// every nw-th chunk.
// ---------------------------------------------------------------------------
struct ObsChunk {
    int o;       // obs index (0 when past the end: a valid row, weighted 0)
    double w;    // rho / rdiag (1 / rdiag with rdiag_wloc), 0 past the end
    double d;    // departure
};

__device__ __forceinline__ ObsChunk stage_chunk(const LekfJob& J, long off, int p, int q0,
                                                int lane, bool wloc, double& p3)
{
    ObsChunk c{0, 0.0, 0.0};
    const int q = q0 + lane;
    if (q < p) {
        c.o = J.loc_idx[off + q];
        const double rho = J.loc_rho[off + q];
        const double rd = J.loc_rdiag ? J.loc_rdiag[off + q] : J.rdiag[c.o];
        c.w = (wloc ? 1.0 : rho) / rd;
        c.d = J.dep[c.o];
        p3 += rho;
    }
    return c;
}

template <int NT>
__device__ __forceinline__ void gram_f64(const LekfJob& J, long off, int p, int k, int lane,
                                         lekf_d4* acc, double& p3, int wv, int nw)
{
    // 4-obs steps whose rows are gathered at once (bounded by the registers
    // the rows and the NT (NT + 1) / 2 accumulator tiles take)
    constexpr int SUB = NT == 1 ? 16 : NT == 2 ? 8 : NT == 3 ? 4 : NT == 4 ? 2 : 1;
    const int lr = lane & 15, kq = lane >> 4;
    const bool wloc = (J.flags & LEKF_F_WLOC) != 0;
    for (int q0 = 64 * wv; q0 < p; q0 += 64 * nw) {
        const ObsChunk ch = stage_chunk(J, off, p, q0, lane, wloc, p3);
#pragma unroll 1
        for (int h = 0; h < 16 && q0 + 4 * h < p; h += SUB) {
            double v[SUB][NT], wt[SUB];
#pragma unroll
            for (int st = 0; st < SUB; ++st) {
                const int src = 4 * (h + st) + kq;
                const int o = __shfl(ch.o, src, 64);
                const double w = __shfl(ch.w, src, 64);
                const double dd = __shfl(ch.d, src, 64);
                wt[st] = w;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int c = 16 * t + lr;
                    double x = 0.0;
                    if (c < k)
                        x = J.hdxf[(long)o * J.y_so + (long)c * J.y_sm];
                    else if (c == k)
                        x = dd;
                    v[st][t] = x;
                }
            }
#pragma unroll
            for (int st = 0; st < SUB; ++st) {
                int a = 0;
#pragma unroll
                for (int I = 0; I < NT; ++I) {
                    const double ai = wt[st] * v[st][I];
#pragma unroll
                    for (int Jt = I; Jt < NT; ++Jt) {
                        acc[a] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, v[st][Jt], acc[a], 0, 0, 0);
                        ++a;
                    }
                }
            }
        }
    }
}

// Mixed form: Y from the FP32 table, products accumulated in FP32 on MFMA.
// FP32: v_mfma_f32_16x16x4f32 (operands rounded to FP32).
// BF16x3: each FP32 operand split x = h + l (h = bf16(x), l = bf16(x - h));
// A*B ~ Ah*Bh + Ah*Bl + Al*Bh on v_mfma_f32_16x16x32_bf16.
// FP32 form: the FP32 accumulators are added into FP64 ones (accd) after each
// chunk of 64 obs, so no FP32 sum runs over more than 64 products and the
// rounding of long local lists (p ~ 10^3) does not build up.  The BF16x3 form
// keeps FP32 accumulation: its two-term split carries about 16 significant
// bits, which dominates its error, and the FP64 accumulators would cost it
// occupancy.
template <int NT>
__device__ __forceinline__ void mix_flush(lekf_f4* acc, lekf_d4* accd)
{
#pragma unroll
    for (int a = 0; a < NT * (NT + 1) / 2; ++a) {
#pragma unroll
        for (int r = 0; r < 4; ++r) accd[a][r] += (double)acc[a][r];
        acc[a] = lekf_f4{0.f, 0.f, 0.f, 0.f};
    }
}

template <int NT>
__device__ __forceinline__ void gram_f32(const LekfJob& J, long off, int p, int k, int lane,
                                         lekf_f4* acc, lekf_d4* accd, double& p3, int wv, int nw)
{
    constexpr int SUB = NT == 1 ? 16 : NT == 2 ? 8 : NT == 3 ? 4 : NT == 4 ? 2 : 1;
    const int lr = lane & 15, kq = lane >> 4;
    const bool wloc = (J.flags & LEKF_F_WLOC) != 0;
    for (int q0 = 64 * wv; q0 < p; q0 += 64 * nw) {
        const ObsChunk ch = stage_chunk(J, off, p, q0, lane, wloc, p3);
        const float wf = (float)ch.w, df = (float)ch.d;  // one dword each across lanes
#pragma unroll 1
        for (int h = 0; h < 16 && q0 + 4 * h < p; h += SUB) {
            float v[SUB][NT], wt[SUB];
#pragma unroll
            for (int st = 0; st < SUB; ++st) {
                const int src = 4 * (h + st) + kq;
                const int o = __shfl(ch.o, src, 64);
                const float w = __shfl(wf, src, 64);
                const float dd = __shfl(df, src, 64);
                wt[st] = w;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int c = 16 * t + lr;
                    float x = 0.0f;
                    if (c < k)
                        x = J.hdxf32[(long)o * k + c];
                    else if (c == k)
                        x = dd;
                    v[st][t] = x;
                }
            }
#pragma unroll
            for (int st = 0; st < SUB; ++st) {
                int a = 0;
#pragma unroll
                for (int I = 0; I < NT; ++I) {
                    const float ai = wt[st] * v[st][I];
#pragma unroll
                    for (int Jt = I; Jt < NT; ++Jt) {
                        acc[a] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, v[st][Jt], acc[a], 0, 0, 0);
                        ++a;
                    }
                }
            }
        }
        mix_flush<NT>(acc, accd);
    }
}

// Split-BF16x3: K = 32 obs per MFMA step, two steps per chunk.  Lane l
// supplies 8 consecutive k values (kk = 8*(l/16) .. +7) of row/col l%16.
template <int NT>
__device__ __forceinline__ void gram_bf16x3(const LekfJob& J, long off, int p, int k, int lane,
                                            lekf_f4* acc, double& p3, int wv, int nw)
{
    const int lr = lane & 15, kq = lane >> 4;
    const bool wloc = (J.flags & LEKF_F_WLOC) != 0;
    for (int q0 = 64 * wv; q0 < p; q0 += 64 * nw) {
        const ObsChunk ch = stage_chunk(J, off, p, q0, lane, wloc, p3);
        // the weight and departure in the precision the products use, so each
        // crosses lanes as one dword
        const float wf = (float)ch.w, df = (float)ch.d;
        for (int h = 0; h < 64 && q0 + h < p; h += 32) {
            lekf_bf8 ah[NT], al[NT], bh[NT], bl[NT];
            float wv8[8], dv8[8];
            int ov[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int src = h + 8 * kq + e;
                ov[e] = __shfl(ch.o, src, 64);
                wv8[e] = __shfl(wf, src, 64);
                dv8[e] = __shfl(df, src, 64);
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const int c = 16 * t + lr;
                float x[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    x[e] = 0.0f;
                    if (c < k)
                        x[e] = J.hdxf32[(long)ov[e] * k + c];
                    else if (c == k)
                        x[e] = dv8[e];
                }
                // two elements per rounding instruction (round to nearest even,
                // as the scalar conversion does)
#pragma unroll
                for (int e = 0; e < 8; e += 2) {
                    const lekf_f2v xv = {x[e], x[e + 1]};
                    const lekf_f2v xa = {wv8[e] * x[e], wv8[e + 1] * x[e + 1]};
                    const lekf_b2v hb = __builtin_convertvector(xv, lekf_b2v);
                    const lekf_b2v ha = __builtin_convertvector(xa, lekf_b2v);
                    const lekf_f2v rb = {x[e] - (float)hb.x, x[e + 1] - (float)hb.y};
                    const lekf_f2v ra = {xa.x - (float)ha.x, xa.y - (float)ha.y};
                    const lekf_b2v lb = __builtin_convertvector(rb, lekf_b2v);
                    const lekf_b2v la = __builtin_convertvector(ra, lekf_b2v);
                    bh[t][e] = hb.x;
                    bh[t][e + 1] = hb.y;
                    bl[t][e] = lb.x;
                    bl[t][e + 1] = lb.y;
                    ah[t][e] = ha.x;
                    ah[t][e + 1] = ha.y;
                    al[t][e] = la.x;
                    al[t][e + 1] = la.y;
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
}

// Write the upper tiles of G into LDS: C (k x k at A, leading dimension lda,
// exactly symmetric), b and sum w d^2; with add, accumulate onto what an
// earlier wave stored.  Only (row <= col) elements are taken so C is
// symmetric bit for bit.  Result-tile layouts (measured on gfx950): the FP64
// 16x16x4 MFMA puts element (row = l/16 + 4r, col = l%16) in register r of
// lane l; the FP32 and BF16 16x16 forms put (row = 4*(l/16) + r, col = l%16)
// there.
template <int NT, bool F64, typename ACC>
__device__ __forceinline__ void store_gram(const ACC* acc, int k, int lda, int lane, double* A,
                                           double* B, double* sumwd2, bool add)
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
                        const double v = add ? A[row * lda + col] + val : val;
                        A[row * lda + col] = v;
                        A[col * lda + row] = v;
                    } else if (col == k) {
                        if (row < k)
                            B[row] = add ? B[row] + val : val;
                        else
                            *sumwd2 = add ? *sumwd2 + val : val;
                    }
                }
            }
            ++a;
        }
    }
}

// Adaptive inflation update (common_letkf.f90:218-239)
__device__ __forceinline__ double adaptive_alpha(double alpha, double p1, double p2, double p3)
{
    const double p4 = (p1 - p3) / p2 - alpha;
    const double qv = (alpha * p2 + p3) / p2;
    const double so = 2.0 / p3 * (qv * qv);
    const double sb2 = 0.04 * 0.04;
    return alpha + sb2 / (so + sb2) * p4;
}

// ===========================================================================
// k <= 32: fused kernel, several points per wave, register Jacobi
// ===========================================================================

struct SolveLds {
    double* A;    // n x lda   C (the Gram), later V (eigenvectors, columns)
    double* B;    // k   (Y^T R^-1 d)
    double* L;    // n   eigenvalues
    double* U;    // k   L^-1 V^T b, later sqrt((k-1)/L)
    double* Wb;   // k   w-bar
    double* X;    // nvmax * kmax  state perturbation rows of the apply
    double* M;    // nvmax         row means
    double* Z;    // nvmax * kmax  X' V diag(U)
    double* D;    // nvmax         X' w-bar
    double* sc;   // 8 scalars: sum w d^2, P2, P3, alpha, k, has obs, status
};

// LDS of the fused kernel for ensemble sizes up to kmax and up to nvmax state
// rows per point: one region per point of the wave, then a region the points'
// final stages share in turn.
//   point region: C (the Gram; the Jacobi reads it and writes V over it), b,
//                 the eigenvalues, 8 scalars
//   shared:       U, w-bar, the apply rows X, their means M, X' V diag(U) and
//                 X' w-bar
struct FusedLayout {
    int kmax, nvmax, n, lda, offB, offL, offS, per, offWb, offX, offM, offZ, offD;
};

__host__ __device__ inline FusedLayout fused_layout(int kmax, int nvmax)
{
    FusedLayout f;
    f.kmax = kmax;
    f.nvmax = nvmax;
    f.n = (kmax + 1) & ~1;
    f.lda = f.n + 1;
    f.offB = f.n * f.lda;
    f.offL = f.offB + kmax;
    f.offS = f.offL + f.n;
    f.per = f.offS + 8;
    f.offWb = kmax;
    f.offX = f.offWb + kmax;
    f.offM = f.offX + nvmax * kmax;
    f.offZ = f.offM + nvmax;
    f.offD = f.offZ + nvmax * kmax;
    return f;
}

__host__ __device__ inline size_t solve_lds_doubles(int kmax, int nvmax, int ppw)
{
    const FusedLayout f = fused_layout(kmax, nvmax);
    return (size_t)ppw * f.per + (size_t)f.offD + (size_t)nvmax;
}

// Views of point region base: C and later V, b, eigenvalues and scalars in the
// point's own region; the final stage's vectors in the shared region.
__device__ inline SolveLds carve_point(double* base, double* shared, const FusedLayout& f)
{
    SolveLds L;
    L.A = base;
    L.B = base + f.offB;
    L.L = base + f.offL;
    L.sc = base + f.offS;
    L.U = shared;
    L.Wb = shared + f.offWb;
    L.X = shared + f.offX;
    L.M = shared + f.offM;
    L.Z = shared + f.offZ;
    L.D = shared + f.offD;
    return L;
}

// Partner of index j in round r of the circle-method tournament on n (even)
// indices: player 0 stays, the others rotate; over the n-1 rounds every pair
// meets once.
__device__ __forceinline__ int rr_partner(int n, int r, int j)
{
    const int m = n - 1;
    if (j == 0) return 1 + (r % m);
    const int x = (j - 1 - (r % m) + m) % m;
    return x == 0 ? 0 : 1 + ((m - x) + r) % m;
}

// Gram stage of one point: C + (k-1)/alpha I into the point region, b, the
// scalars of the inflation update, and the point's k and obs flag in sc.
template <int NT, int MODE>
__device__ __forceinline__ void fused_gram(const LekfJob& J, int b, SolveLds& L, int lane)
{
    const int k = J.k;
    const int n = (k + 1) & ~1, lda = n + 1;
    const long off = J.loc_off[b];
    const int p = J.loc_cnt ? J.loc_cnt[b] : (int)(J.loc_off[b + 1] - off);
    double alpha = J.infl_in ? J.infl_in[b] : J.infl_fixed;
    if (p > 0) {
        double p3 = 0.0;
        if (MODE == 0) {
            lekf_d4 acc[NT * (NT + 1) / 2];
#pragma unroll
            for (int a = 0; a < NT * (NT + 1) / 2; ++a) acc[a] = lekf_d4{0.0, 0.0, 0.0, 0.0};
            gram_f64<NT>(J, off, p, k, lane, acc, p3, 0, 1);
            store_gram<NT, true>(acc, k, lda, lane, L.A, L.B, &L.sc[0], false);
        } else {
            lekf_f4 acc[NT * (NT + 1) / 2];
#pragma unroll
            for (int a = 0; a < NT * (NT + 1) / 2; ++a) acc[a] = lekf_f4{0.f, 0.f, 0.f, 0.f};
            if (MODE == 1) {
                lekf_d4 accd[NT * (NT + 1) / 2];
#pragma unroll
                for (int a = 0; a < NT * (NT + 1) / 2; ++a) accd[a] = lekf_d4{0.0, 0.0, 0.0, 0.0};
                gram_f32<NT>(J, off, p, k, lane, acc, accd, p3, 0, 1);
                store_gram<NT, false>(accd, k, lda, lane, L.A, L.B, &L.sc[0], false);
            } else {
                gram_bf16x3<NT>(J, off, p, k, lane, acc, p3, 0, 1);
                store_gram<NT, false>(acc, k, lda, lane, L.A, L.B, &L.sc[0], false);
            }
        }
        p3 = wave_sum(p3);
        __syncthreads();
        if (lane == 0) {
            double tr = 0.0;
            for (int i = 0; i < k; ++i) tr += L.A[i * lda + i];
            L.sc[1] = tr / (double)(k - 1);  // P2
            L.sc[2] = p3;                    // P3
            if ((J.flags & LEKF_F_MINFL) && J.minfl > 0.0 && alpha < J.minfl) alpha = J.minfl;
            const double rho = 1.0 / alpha;
            for (int i = 0; i < k; ++i) L.A[i * lda + i] += (double)(k - 1) * rho;
        }
    }
    if (lane == 0) {
        L.sc[3] = alpha;
        L.sc[4] = (double)k;
        L.sc[5] = p > 0 ? 1.0 : 0.0;
        L.sc[6] = 0.0;
    }
}

// Eigendecomposition of PPW points at once, for KR <= 32.  Lane group g (lanes
// g*KR .. g*KR+KR-1) holds point g, lane j of a group column j of its matrix.
//   1. Cholesky C = R^T R, column j in lane j: step s takes R[s][s] from its
//      lane by a shuffle and row s of R through the point region, where R
//      stays.
//   2. One-sided (Hestenes) cyclic Jacobi on the columns of R, tournament
//      ordering: each round a lane fetches its partner's column, forms
//      [alpha gamma; gamma beta] (the 2x2 Gram of the pair) and the rotation
//      that zeroes gamma, and applies it to its column.  At convergence
//      R J = U S with orthogonal columns, so C = R^T R = J S^2 J^T: the
//      eigenvalues are the squared column norms.
//   3. V = J = R^-1 (R J), one back substitution per lane with R read from
//      the point region.
// Only the columns of R travel between lanes (not V as well), and the
// eigenvectors come from a triangular solve with R, whose condition number is
// the square root of C's.  Rows and columns >= k are zero and never rotate;
// so does a point with no obs, and a point that has converged while others of
// the wave still rotate, so a point's result does not depend on the other
// points of its wave.  Writes the eigenvalues and V over C in each point
// region, and LEKF_PT_NOCONV into sc[6] of a point still rotating at the
// sweep limit.
template <int KR, int PPW>
__device__ __forceinline__ void jacobi_multi(double* smem, const FusedLayout& f, int lane)
{
    const int g = lane / KR, j = lane - g * KR;
    const bool in = g < PPW;
    double* base = smem + (size_t)(in ? g : 0) * f.per;
    const int k = in ? (int)base[f.offS + 4] : 0;
    const int kk = (in && base[f.offS + 5] != 0.0) ? k : 0;
    const int lda = ((k + 1) & ~1) + 1;
    const int gb = in ? g * KR : lane;  // lane of column 0 of this group (idle lanes: themselves)
    double x[KR];
#pragma unroll
    for (int i = 0; i < KR; ++i) x[i] = (j < kk && i < kk) ? base[i * lda + j] : 0.0;
    // 1. Cholesky, right-looking: at step s lane j (j >= s) forms R[s][j] in
    // x[s] and writes it to row s of the point region (which C no longer
    // needs: every lane holds its column), then takes R[s][i] R[s][j] off its
    // entries i in (s, j] with row s read back from the region
    __syncthreads();  // every lane has read its column of C
#pragma unroll
    for (int s0 = 0; s0 < KR; ++s0) {
        const double d = __shfl(x[s0], in ? gb + s0 : lane, 64);
        const double rss = d > 0.0 ? sqrt(d) : 0.0;
        const double inv = rss > 0.0 ? 1.0 / rss : 0.0;
        x[s0] = j > s0 ? x[s0] * inv : (j == s0 ? rss : 0.0);
        if (in && j < kk && s0 < kk) base[s0 * lda + j] = x[s0];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < KR; ++i)  // constant trip count: unrolls fully
            if (i > s0 && j >= i && i < kk) x[i] = x[i] - base[s0 * lda + i] * x[s0];
    }
    // 2. Jacobi on the columns of R
    double alpha = 0.0;
#pragma unroll
    for (int i = 0; i < KR; ++i) alpha += x[i] * x[i];
    const unsigned long long gmask = (KR < 64 ? (1ull << KR) : 0ull) - 1ull;
    int sweep = 0;
    bool rotated = false;
    bool done = false;  // this lane's point has converged
    for (; sweep < LEKF_JACOBI_MAX_SWEEPS; ++sweep) {
        rotated = false;
        bool big = false;  // a pair further than 1e-8 from orthogonal was rotated
        for (int r = 0; r < KR - 1; ++r) {
            const int pj = rr_partner(KR, r, j);
            const int src = in ? gb + pj : lane;
            double y[KR];
#pragma unroll
            for (int i = 0; i < KR; ++i) y[i] = __shfl(x[i], src, 64);
            double gamma = 0.0;
#pragma unroll
            for (int i = 0; i < KR; ++i) gamma += x[i] * y[i];
            const double beta = __shfl(alpha, src, 64);
            const bool lo = j < pj;
            const double app = lo ? alpha : beta, aqq = lo ? beta : alpha;
            double c = 1.0, s = 0.0, t = 0.0;
            const double g2 = done ? 0.0 : gamma * gamma;
            const double aa = app * aqq;
            const bool rot = g2 > (LEKF_JACOBI_TOL * LEKF_JACOBI_TOL) * aa;
            big |= g2 > 1.0e-16 * aa;
            if (rot) {
                const double th = (aqq - app) * rcp_nr(2.0 * gamma);
                const double q = fma(th, th, 1.0);
                const double at = fabs(th) + q * rsq_nr(q);
                t = th >= 0.0 ? rcp_nr(at) : -rcp_nr(at);
                c = rsq_nr(fma(t, t, 1.0));
                s = t * c;
            }
            const double sg = lo ? -s : s;
#pragma unroll
            for (int i = 0; i < KR; ++i) x[i] = c * x[i] + sg * y[i];
            if (rot) {
                alpha = lo ? app - t * gamma : aqq + t * gamma;
                rotated = true;
            }
        }
        // Quadratic convergence: a sweep that started with every pair of a
        // point within 1e-8 of orthogonal leaves them within about 1e-16, so
        // the point stops instead of running a sweep that would only confirm
        // it.  The test is per point, so a point's sweeps do not depend on
        // the other points of its wave.
        const unsigned long long bigm = __ballot(big);
        if (in && ((bigm >> (g * KR)) & gmask) == 0ull) done = true;
        rotated = rotated && !done;
        if (__ballot(rotated) == 0ull) break;
        // refresh the column norms once per sweep
        alpha = 0.0;
#pragma unroll
        for (int i = 0; i < KR; ++i) alpha += x[i] * x[i];
    }
    const unsigned long long still = __ballot(in && rotated);
    double nrm = 0.0;
#pragma unroll
    for (int i = 0; i < KR; ++i) nrm += x[i] * x[i];
    // 3. V = R^-1 (R J): back substitution with R from the point region, in
    // place (x[l], l > ii, already holds V's entry l when row ii is solved)
    __syncthreads();
#pragma unroll
    for (int ii = KR - 1; ii >= 0; --ii) {
        double sacc = x[ii];
#pragma unroll
        for (int l = 0; l < KR; ++l)  // constant trip count: unrolls fully
            if (l > ii && l < kk) sacc = sacc - base[ii * lda + l] * x[l];
        const double rii = ii < kk ? base[ii * lda + ii] : 0.0;
        x[ii] = rii > 0.0 ? sacc / rii : 0.0;
    }
    __syncthreads();  // every lane has read R before V goes over it
    if (in && j < kk) {
        base[f.offL + j] = nrm;
#pragma unroll
        for (int i = 0; i < KR; ++i)
            if (i < kk) base[i * lda + j] = x[i];
    }
    if (in && j == 0 && sweep == LEKF_JACOBI_MAX_SWEEPS && ((still >> (g * KR)) & gmask) != 0ull)
        base[f.offS + 6] = (double)LEKF_PT_NOCONV;
}

// Final stage of one point: eigenvalue test, w-bar, the inflation update,
// the optional outputs and the fused apply.  With T = V diag(U) V^T + w-bar 1^T
// (U = sqrt((k-1)/lambda)), the apply row is x-mean + (X' V) diag(U) V^T +
// (X' . w-bar) 1^T, so T itself is formed only when the caller asks for it.
__device__ __forceinline__ void fused_finish(const LekfJob& J, int b, SolveLds& L, int lane)
{
    const int k = J.k;
    const int n = (k + 1) & ~1, lda = n + 1;
    const double* V = L.A;
    const double alpha = L.sc[3];
    const bool has_obs = L.sc[5] != 0.0;
    const bool transm_mode = (J.flags & LEKF_F_TRANSM) != 0;
    int status = (int)L.sc[6];
    if (has_obs) {
        // eigenvalues; reference truncation test (common_mtx.f90:66-74)
        double lmax = 0.0, lmin = 1.0e308;
        for (int j = 0; j < k; ++j) {
            const double lj = L.L[j];
            lmax = lj > lmax ? lj : lmax;
            lmin = lj < lmin ? lj : lmin;
        }
        if (!(lmax > 0.0) || lmin < fabs(lmax) * 1.4901161193847656e-08) status |= LEKF_PT_RANKDEF;
        // u = L^-1 V^T b ; w-bar = V u
        for (int j = lane; j < k; j += LEKF_WAVE) {
            double s = 0.0;
            for (int i = 0; i < k; ++i) s += V[i * lda + j] * L.B[i];
            L.U[j] = s / L.L[j];
        }
        __syncthreads();
        for (int i = lane; i < k; i += LEKF_WAVE) {
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += V[i * lda + j] * L.U[j];
            L.Wb[i] = s;
        }
        // U <- sqrt((k-1)/lambda)
        __syncthreads();
        for (int j = lane; j < k; j += LEKF_WAVE) L.U[j] = sqrt((double)(k - 1) / L.L[j]);
        __syncthreads();
        if (J.trans) {
            double* tr = J.trans + (size_t)b * k * k;
            for (int w = lane; w < k * k; w += LEKF_WAVE) {
                const int m = w / k, i = w - m * k;  // column-major: coalesced over i
                double s = 0.0;
                for (int l = 0; l < k; ++l) s += V[i * lda + l] * L.U[l] * V[m * lda + l];
                if (!transm_mode) s += L.Wb[i];
                tr[i + (size_t)m * k] = s;
            }
        }
        if (J.pao) {
            double* pa = J.pao + (size_t)b * k * k;
            for (int w = lane; w < k * k; w += LEKF_WAVE) {
                const int i = w / k, m = w - i * k;
                double s = 0.0;
                for (int l = 0; l < k; ++l) s += V[i * lda + l] / L.L[l] * V[m * lda + l];
                pa[i + (size_t)m * k] = s;
            }
        }
        if (J.transm)
            for (int i = lane; i < k; i += LEKF_WAVE) J.transm[(size_t)b * k + i] = L.Wb[i];
        if (lane == 0 && J.infl_out)
            J.infl_out[b] = (J.flags & LEKF_F_FIXED_INFL)
                                ? J.infl_fixed
                                : adaptive_alpha(alpha, L.sc[0], L.sc[1], L.sc[2]);
    } else {
        // no local obs (common_letkf.f90:82-96): T = sqrt(alpha) I
        const double sa = sqrt(alpha);
        if (J.trans) {
            double* tr = J.trans + (size_t)b * k * k;
            for (int w = lane; w < k * k; w += LEKF_WAVE) {
                const int m = w / k, i = w - m * k;
                tr[i + (size_t)m * k] = (i == m) ? sa : 0.0;
            }
        }
        if (J.pao) {
            double* pa = J.pao + (size_t)b * k * k;
            for (int w = lane; w < k * k; w += LEKF_WAVE) {
                const int i = w / k, m = w - i * k;
                pa[i + (size_t)m * k] = (i == m) ? alpha / (double)(k - 1) : 0.0;
            }
        }
        if (J.transm)
            for (int i = lane; i < k; i += LEKF_WAVE) J.transm[(size_t)b * k + i] = 0.0;
        if (lane == 0 && J.infl_out)
            J.infl_out[b] = (J.flags & LEKF_F_FIXED_INFL) ? J.infl_fixed : alpha;
    }
    if (lane == 0 && J.status) J.status[b] = status;
    if (!(J.flags & LEKF_F_APPLY)) return;
    // fused apply: rows of the point (3-D at every level, 2-D at level 0)
    const int lev = b / J.nij, ij = b - lev * J.nij;
    const int n3 = J.nv3d, n2 = (lev == 0) ? J.nv2d : 0;
    const int nrow = n3 + n2;
    const uint32_t mask = J.upd_mask ? J.upd_mask[lev] : 0xffffffffu;
    for (int w = lane; w < nrow * k; w += LEKF_WAVE) {
        const int r = w / k, m = w - r * k;
        const double* src = (r < n3) ? J.x3b + ((size_t)b * n3 + r) * k
                                     : J.x2b + ((size_t)ij * J.nv2d + (r - n3)) * k;
        L.X[r * k + m] = src[m];
    }
    __syncthreads();
    // ensemble mean in member order (speedy/common/common_speedy.f90:335-370)
    for (int r = lane; r < nrow; r += LEKF_WAVE) {
        double s = L.X[r * k];
        for (int m = 1; m < k; ++m) s += L.X[r * k + m];
        L.M[r] = s / (double)k;
    }
    __syncthreads();
    for (int w = lane; w < nrow * k; w += LEKF_WAVE) L.X[w] -= L.M[w / k];
    __syncthreads();
    if (has_obs) {
        // Z = X' V diag(U) and D = X' w-bar, row by row
        for (int w = lane; w < nrow * k; w += LEKF_WAVE) {
            const int r = w / k, l = w - r * k;
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += L.X[r * k + j] * V[j * lda + l];
            L.Z[r * k + l] = s * L.U[l];
        }
        for (int r = lane; r < nrow; r += LEKF_WAVE) {
            double d = 0.0;
            if (!transm_mode)
                for (int j = 0; j < k; ++j) d += L.X[r * k + j] * L.Wb[j];
            L.D[r] = d;
        }
        __syncthreads();
    }
    const double sa = has_obs ? 0.0 : sqrt(alpha);
    for (int w = lane; w < nrow * k; w += LEKF_WAVE) {
        const int r = w / k, m = w - r * k;
        const bool upd = (r >= n3) || ((mask >> r) & 1u);
        double a = L.M[r];
        if (!upd) {
            a += L.X[r * k + m];
        } else if (has_obs) {
            double s = 0.0;
            for (int l = 0; l < k; ++l) s += L.Z[r * k + l] * V[m * lda + l];
            a += s + L.D[r];
        } else {
            a += sa * L.X[r * k + m];
        }
        double* dst = (r < n3) ? J.x3a + ((size_t)b * n3 + r) * k
                               : J.x2a + ((size_t)ij * J.nv2d + (r - n3)) * k;
        dst[m] = a;
    }
}

// NT: 16-wide Gram tiles; KR: rows of the register Jacobi (k <= KR <= 32);
// MODE 0: FP64 Gram, 1: FP32, 2: split-BF16x3; PPW points per wave (lane
// groups of KR for the Jacobi).  Points gp = blockIdx.x * PPW + g of the flat
// point range [0, npt); every job of the launch has k <= kmax.
template <int NT, int KR, int MODE, int PPW>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(KR <= 20 ? 3 : 2)))
void lekf_solve_kernel(const LekfJob* __restrict__ jobs, int njobs, long npt, int kmax, int nvmax)
{
    extern __shared__ double smem[];
    const int lane = threadIdx.x;
    const FusedLayout f = fused_layout(kmax, nvmax);
    double* const shared = smem + (size_t)PPW * f.per;
    const long gp0 = (long)blockIdx.x * PPW;

    for (int g = 0; g < PPW; ++g) {
        double* base = smem + (size_t)g * f.per;
        const long gp = gp0 + g;
        if (gp >= npt) {
            if (lane == 0) {
                base[f.offS + 4] = 0.0;
                base[f.offS + 5] = 0.0;
            }
            continue;
        }
        const LekfJob& J = jobs[lekf_job_of(jobs, njobs, gp)];
        const int b = (int)(gp - J.pt_begin + J.pt0);
        SolveLds L = carve_point(base, shared, f);
        fused_gram<NT, MODE>(J, b, L, lane);
    }
    __syncthreads();
    jacobi_multi<KR, PPW>(smem, f, lane);
    __syncthreads();
    for (int g = 0; g < PPW; ++g) {
        const long gp = gp0 + g;
        if (gp >= npt) break;
        const LekfJob& J = jobs[lekf_job_of(jobs, njobs, gp)];
        const int b = (int)(gp - J.pt_begin + J.pt0);
        SolveLds L = carve_point(smem + (size_t)g * f.per, shared, f);
        fused_finish(J, b, L, lane);
        __syncthreads();
    }
}

// ===========================================================================
// 32 < k <= LEKF_KMAX: Gram kernel + eigen/assembly/apply kernel
// ===========================================================================

// workspace record of one point: C (k x k, row-major), b (k), sum w d^2, P3
__host__ __device__ inline long gram_stride(int kmax)
{
    return ((long)kmax * kmax + kmax + 2 + 1) & ~1L;
}

// split-record flag: the mixed-precision pass found C ill-conditioned and put
// the point on the redo list, whose points a second pass solves again from an
// FP64 Gram after the launch's last sub-batch
constexpr int kSplitRedo = 16;

// Waves per SIMD asked of the Gram kernel: the FP64 and FP32 forms with
// 4 x 4 tiles (49 <= k <= 64) take 214 and 188 registers, two waves per
// SIMD, unless held to three's budget (which they fit without spilling)
constexpr int gram_waves_per_simd(int nt, int mode)
{
    return (nt == 4 && mode != 2) ? 3 : (nt == 3 && mode != 2) ? 4 : 1;
}

// W waves per point; wave w takes every W-th chunk of 64 local obs, and the
// partial Grams are summed in LDS in wave order (deterministic for a given W).
template <int NT, int MODE, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(gram_waves_per_simd(NT, MODE))))
void lekf_gram_kernel(
    const LekfJob* __restrict__ jobs, int njobs, long gp0, double* __restrict__ ws, long gstride,
    const int* __restrict__ rlist)
{
    extern __shared__ double smem[];
    const long slot = blockIdx.x;
    long gp = gp0 + slot;
    if (rlist) {  // FP64 pass of a split launch: slot i holds the redo list's i-th point
        if (slot >= rlist[0]) return;
        gp = rlist[1 + slot];
    }
    const int tid = threadIdx.x;
    const int lane = tid & (LEKF_WAVE - 1), wv = tid / LEKF_WAVE;
    const int ji = lekf_job_of(jobs, njobs, gp);
    const LekfJob& J = jobs[ji];
    const int b = (int)(gp - J.pt_begin + J.pt0);
    const int k = J.k, lda = k + 1;
    const long off = J.loc_off[b];
    const int p = J.loc_cnt ? J.loc_cnt[b] : (int)(J.loc_off[b + 1] - off);
    if (p <= 0) return;
    double* A = smem;
    double* B = A + (size_t)k * lda;
    double* sc = B + k;  // [0] sum w d^2, [4 + w] P3 of wave w (4 + W slots)
    double p3 = 0.0;
    if (MODE == 0) {
        lekf_d4 acc[NT * (NT + 1) / 2];
#pragma unroll
        for (int a = 0; a < NT * (NT + 1) / 2; ++a) acc[a] = lekf_d4{0.0, 0.0, 0.0, 0.0};
        gram_f64<NT>(J, off, p, k, lane, acc, p3, wv, W);
        for (int w = 0; w < W; ++w) {
            if (wv == w) store_gram<NT, true>(acc, k, lda, lane, A, B, &sc[0], w > 0);
            __syncthreads();
        }
    } else {
        lekf_f4 acc[NT * (NT + 1) / 2];
#pragma unroll
        for (int a = 0; a < NT * (NT + 1) / 2; ++a) acc[a] = lekf_f4{0.f, 0.f, 0.f, 0.f};
        if (MODE == 1) {
            lekf_d4 accd[NT * (NT + 1) / 2];
#pragma unroll
            for (int a = 0; a < NT * (NT + 1) / 2; ++a) accd[a] = lekf_d4{0.0, 0.0, 0.0, 0.0};
            gram_f32<NT>(J, off, p, k, lane, acc, accd, p3, wv, W);
            for (int w = 0; w < W; ++w) {
                if (wv == w) store_gram<NT, false>(accd, k, lda, lane, A, B, &sc[0], w > 0);
                __syncthreads();
            }
        } else {
            gram_bf16x3<NT>(J, off, p, k, lane, acc, p3, wv, W);
            for (int w = 0; w < W; ++w) {
                if (wv == w) store_gram<NT, false>(acc, k, lda, lane, A, B, &sc[0], w > 0);
                __syncthreads();
            }
        }
    }
    p3 = wave_sum(p3);
    if (lane == 0) sc[4 + wv] = p3;
    __syncthreads();
    double* out = ws + slot * gstride;
    for (int t = tid; t < k * k; t += 64 * W) {
        const int i = t / k, j = t - i * k;
        out[t] = A[i * lda + j];
    }
    for (int i = tid; i < k; i += 64 * W) out[(size_t)k * k + i] = B[i];
    if (tid == 0) {
        double s = sc[4];
        for (int w = 1; w < W; ++w) s += sc[4 + w];
        out[(size_t)k * k + k] = sc[0];
        out[(size_t)k * k + k + 1] = s;
    }
}

// ---------------------------------------------------------------------------
// Block Gram for LEKF_KSPLIT < k <= LEKF_KMAX: the augmented (k+1)^2 Gram is
// cut into 64 x 64 blocks; one wave computes one upper block (bi <= bj) of one
// point over all of its obs, as NB x NB MFMA tiles, and writes it to the
// point's workspace record (both triangles of C, b and sum w d^2).
// ---------------------------------------------------------------------------
constexpr int kBlkNB = 4;  // 16 x 16 tiles per block side

__host__ __device__ inline int gram_nblk(int kmax)
{
    const int nb = (kmax + 1 + 16 * kBlkNB - 1) / (16 * kBlkNB);
    return nb * (nb + 1) / 2;
}

// block pair (bi, bj), bi <= bj, of upper-block index t (row-major over the
// upper triangle)
__device__ __forceinline__ void blk_pair(int t, int nb, int& bi, int& bj)
{
    int i = 0;
    while (t >= nb - i) {
        t -= nb - i;
        ++i;
    }
    bi = i;
    bj = i + t;
}

__device__ __forceinline__ void blk_store(double* out, int k, int grow, int gcol, double val)
{
    if (grow > gcol || gcol > k) return;
    if (gcol < k) {
        out[(size_t)grow * k + gcol] = val;
        out[(size_t)gcol * k + grow] = val;
    } else if (grow < k) {
        out[(size_t)k * k + grow] = val;
    } else {
        out[(size_t)k * k + k] = val;
    }
}

// Large-k points with fewer local obs than members (0 < p < k) and no
// trans/pao output are solved in the dual form: the p x p Gram
// G = B B^T with B = diag(sqrt(w)) Y (one row per local obs), whose
// eigenpairs give those of C = (k-1)/alpha I + B^T B.
__device__ __forceinline__ bool big_dual(const LekfJob& J, int p, int k)
{
    return p > 0 && p < k && !J.trans && !J.pao;
}

// One upper 64 x 64 block of G (p x p, row-major at out) over the k members.
template <int MODE>
__device__ void gram_dual_blk(const LekfJob& J, long off, int p, int k, int tblk, double* out,
                              int lane)
{
    constexpr int NB = kBlkNB;
    const int nbp = (p + 16 * NB - 1) / (16 * NB);
    if (tblk >= nbp * (nbp + 1) / 2) return;
    int bi, bj;
    blk_pair(tblk, nbp, bi, bj);
    const int cI = 16 * NB * bi, cJ = 16 * NB * bj;
    const int lr = lane & 15, kq = lane >> 4;
    const bool wloc = (J.flags & LEKF_F_WLOC) != 0;
    long orow[NB], ocol[NB];
    double swr[NB], swc[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) {
        const int i = cI + 16 * t + lr, j = cJ + 16 * t + lr;
        orow[t] = ocol[t] = -1;
        swr[t] = swc[t] = 0.0;
        if (i < p) {
            const long o = J.loc_idx[off + i];
            const double rd = J.loc_rdiag ? J.loc_rdiag[off + i] : J.rdiag[o];
            swr[t] = sqrt((wloc ? 1.0 : J.loc_rho[off + i]) / rd);
            orow[t] = o;
        }
        if (j < p) {
            const long o = J.loc_idx[off + j];
            const double rd = J.loc_rdiag ? J.loc_rdiag[off + j] : J.rdiag[o];
            swc[t] = sqrt((wloc ? 1.0 : J.loc_rho[off + j]) / rd);
            ocol[t] = o;
        }
    }
    if (MODE == 0) {
        lekf_d4 acc[NB][NB];
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int Jt = 0; Jt < NB; ++Jt) acc[I][Jt] = lekf_d4{0.0, 0.0, 0.0, 0.0};
        for (int m0 = 0; m0 < k; m0 += 4) {
            const int m = m0 + kq;
            double xr[NB], xc[NB];
#pragma unroll
            for (int t = 0; t < NB; ++t) {
                xr[t] = (orow[t] >= 0 && m < k)
                            ? swr[t] * J.hdxf[orow[t] * J.y_so + (long)m * J.y_sm] : 0.0;
                xc[t] = (ocol[t] >= 0 && m < k)
                            ? swc[t] * J.hdxf[ocol[t] * J.y_so + (long)m * J.y_sm] : 0.0;
            }
#pragma unroll
            for (int I = 0; I < NB; ++I)
#pragma unroll
                for (int Jt = 0; Jt < NB; ++Jt)
                    acc[I][Jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[I], xc[Jt], acc[I][Jt], 0, 0, 0);
        }
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int Jt = 0; Jt < NB; ++Jt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = cI + 16 * I + kq + 4 * r, col = cJ + 16 * Jt + lr;
                    if (row <= col && col < p) {
                        out[(size_t)row * p + col] = acc[I][Jt][r];
                        out[(size_t)col * p + row] = acc[I][Jt][r];
                    }
                }
    } else if (MODE == 1) {
        lekf_f4 acc[NB][NB];
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int Jt = 0; Jt < NB; ++Jt) acc[I][Jt] = lekf_f4{0.f, 0.f, 0.f, 0.f};
        float fr[NB], fc[NB];
#pragma unroll
        for (int t = 0; t < NB; ++t) {
            fr[t] = (float)swr[t];
            fc[t] = (float)swc[t];
        }
        for (int m0 = 0; m0 < k; m0 += 4) {
            const int m = m0 + kq;
            float xr[NB], xc[NB];
#pragma unroll
            for (int t = 0; t < NB; ++t) {
                xr[t] = (orow[t] >= 0 && m < k) ? fr[t] * J.hdxf32[orow[t] * (long)k + m] : 0.0f;
                xc[t] = (ocol[t] >= 0 && m < k) ? fc[t] * J.hdxf32[ocol[t] * (long)k + m] : 0.0f;
            }
#pragma unroll
            for (int I = 0; I < NB; ++I)
#pragma unroll
                for (int Jt = 0; Jt < NB; ++Jt)
                    acc[I][Jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[I], xc[Jt], acc[I][Jt], 0, 0, 0);
        }
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int Jt = 0; Jt < NB; ++Jt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = cI + 16 * I + 4 * kq + r, col = cJ + 16 * Jt + lr;
                    if (row <= col && col < p) {
                        out[(size_t)row * p + col] = (double)acc[I][Jt][r];
                        out[(size_t)col * p + row] = (double)acc[I][Jt][r];
                    }
                }
    } else {
        lekf_f4 acc[NB][NB];
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int Jt = 0; Jt < NB; ++Jt) acc[I][Jt] = lekf_f4{0.f, 0.f, 0.f, 0.f};
        float fr[NB], fc[NB];
#pragma unroll
        for (int t = 0; t < NB; ++t) {
            fr[t] = (float)swr[t];
            fc[t] = (float)swc[t];
        }
        for (int m0 = 0; m0 < k; m0 += 32) {
            lekf_bf8 ah[NB], al[NB], bh[NB], bl[NB];
#pragma unroll
            for (int t = 0; t < NB; ++t) {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int m = m0 + 8 * kq + e;
                    const float a1 = (orow[t] >= 0 && m < k) ? fr[t] * J.hdxf32[orow[t] * (long)k + m] : 0.0f;
                    const float a2 = (ocol[t] >= 0 && m < k) ? fc[t] * J.hdxf32[ocol[t] * (long)k + m] : 0.0f;
                    const __bf16 ah1 = (__bf16)a1;
                    const __bf16 al1 = (__bf16)(a1 - (float)ah1);
                    const __bf16 bh1 = (__bf16)a2;
                    const __bf16 bl1 = (__bf16)(a2 - (float)bh1);
                    ah[t][e] = ah1;
                    al[t][e] = al1;
                    bh[t][e] = bh1;
                    bl[t][e] = bl1;
                }
            }
#pragma unroll
            for (int I = 0; I < NB; ++I)
#pragma unroll
                for (int Jt = 0; Jt < NB; ++Jt) {
                    acc[I][Jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[I], bh[Jt], acc[I][Jt], 0, 0, 0);
                    acc[I][Jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[I], bl[Jt], acc[I][Jt], 0, 0, 0);
                    acc[I][Jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[I], bh[Jt], acc[I][Jt], 0, 0, 0);
                }
        }
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int Jt = 0; Jt < NB; ++Jt)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = cI + 16 * I + 4 * kq + r, col = cJ + 16 * Jt + lr;
                    if (row <= col && col < p) {
                        out[(size_t)row * p + col] = (double)acc[I][Jt][r];
                        out[(size_t)col * p + row] = (double)acc[I][Jt][r];
                    }
                }
    }
}

// Block tblk (upper-triangle order) of point b's Gram into its workspace
// record out: [C | b | sum w d^2 | P3] (64 x 64 blocks of the augmented
// (k+1)^2 product), or a block of G in the dual form.
template <int MODE>
__device__ void gram_blk_point(const LekfJob& J, int b, int tblk, double* out, int lane)
{
    constexpr int NB = kBlkNB;
    const int k = J.k;
    const int nbk = (k + 1 + 16 * NB - 1) / (16 * NB);
    if (tblk >= nbk * (nbk + 1) / 2) return;  // block of a larger k of the launch
    int bi, bj;
    blk_pair(tblk, nbk, bi, bj);
    const int cI = 16 * NB * bi, cJ = 16 * NB * bj;
    const long off = J.loc_off[b];
    const int p = J.loc_cnt ? J.loc_cnt[b] : (int)(J.loc_off[b + 1] - off);
    if (p <= 0) return;
    if (big_dual(J, p, k)) {
        gram_dual_blk<MODE>(J, off, p, k, tblk, out, lane);
        return;
    }
    const int lr = lane & 15, kq = lane >> 4;
    const bool wloc = (J.flags & LEKF_F_WLOC) != 0;
    double p3 = 0.0;
    if (MODE == 0) {
        lekf_d4 acc[NB][NB];
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int Jt = 0; Jt < NB; ++Jt) acc[I][Jt] = lekf_d4{0.0, 0.0, 0.0, 0.0};
        for (int q0 = 0; q0 < p; q0 += 4) {
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
            double xr[NB], xc[NB];
#pragma unroll
            for (int t = 0; t < NB; ++t) {
                const int c1 = cI + 16 * t + lr, c2 = cJ + 16 * t + lr;
                double a1 = 0.0, a2 = 0.0;
                if (valid) {
                    a1 = c1 < k ? J.hdxf[o * J.y_so + (long)c1 * J.y_sm] : (c1 == k ? dd : 0.0);
                    a2 = c2 < k ? J.hdxf[o * J.y_so + (long)c2 * J.y_sm] : (c2 == k ? dd : 0.0);
                }
                xr[t] = w * a1;
                xc[t] = a2;
            }
#pragma unroll
            for (int I = 0; I < NB; ++I)
#pragma unroll
                for (int Jt = 0; Jt < NB; ++Jt)
                    acc[I][Jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[I], xc[Jt], acc[I][Jt], 0, 0, 0);
        }
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int Jt = 0; Jt < NB; ++Jt)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    blk_store(out, k, cI + 16 * I + kq + 4 * r, cJ + 16 * Jt + lr, acc[I][Jt][r]);
    } else if (MODE == 1) {
        lekf_f4 acc[NB][NB];
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int Jt = 0; Jt < NB; ++Jt) acc[I][Jt] = lekf_f4{0.f, 0.f, 0.f, 0.f};
        for (int q0 = 0; q0 < p; q0 += 4) {
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
            float xr[NB], xc[NB];
#pragma unroll
            for (int t = 0; t < NB; ++t) {
                const int c1 = cI + 16 * t + lr, c2 = cJ + 16 * t + lr;
                float a1 = 0.0f, a2 = 0.0f;
                if (valid) {
                    a1 = c1 < k ? J.hdxf32[o * (long)k + c1] : (c1 == k ? dd : 0.0f);
                    a2 = c2 < k ? J.hdxf32[o * (long)k + c2] : (c2 == k ? dd : 0.0f);
                }
                xr[t] = w * a1;
                xc[t] = a2;
            }
#pragma unroll
            for (int I = 0; I < NB; ++I)
#pragma unroll
                for (int Jt = 0; Jt < NB; ++Jt)
                    acc[I][Jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[I], xc[Jt], acc[I][Jt], 0, 0, 0);
        }
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int Jt = 0; Jt < NB; ++Jt)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    blk_store(out, k, cI + 16 * I + 4 * kq + r, cJ + 16 * Jt + lr,
                              (double)acc[I][Jt][r]);
    } else {
        lekf_f4 acc[NB][NB];
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int Jt = 0; Jt < NB; ++Jt) acc[I][Jt] = lekf_f4{0.f, 0.f, 0.f, 0.f};
        for (int q0 = 0; q0 < p; q0 += 32) {
            float wv8[8], dv[8];
            long ov[8];
            bool vv[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int q = q0 + 8 * kq + e;
                vv[e] = q < p;
                wv8[e] = 0.0f;
                dv[e] = 0.0f;
                ov[e] = 0;
                if (vv[e]) {
                    const long o = J.loc_idx[off + q];
                    const double rho = J.loc_rho[off + q];
                    const double rd = J.loc_rdiag ? J.loc_rdiag[off + q] : J.rdiag[o];
                    wv8[e] = (float)((wloc ? 1.0 : rho) / rd);
                    dv[e] = (float)J.dep[o];
                    ov[e] = o;
                    if (lr == 0) p3 += rho;
                }
            }
            lekf_bf8 ah[NB], al[NB], bh[NB], bl[NB];
#pragma unroll
            for (int t = 0; t < NB; ++t) {
                const int c1 = cI + 16 * t + lr, c2 = cJ + 16 * t + lr;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    float a1 = 0.0f, a2 = 0.0f;
                    if (vv[e]) {
                        a1 = c1 < k ? J.hdxf32[ov[e] * (long)k + c1] : (c1 == k ? dv[e] : 0.0f);
                        a2 = c2 < k ? J.hdxf32[ov[e] * (long)k + c2] : (c2 == k ? dv[e] : 0.0f);
                    }
                    const float xa = wv8[e] * a1;
                    const __bf16 ah1 = (__bf16)xa;
                    const __bf16 al1 = (__bf16)(xa - (float)ah1);
                    const __bf16 bh1 = (__bf16)a2;
                    const __bf16 bl1 = (__bf16)(a2 - (float)bh1);
                    ah[t][e] = ah1;
                    al[t][e] = al1;
                    bh[t][e] = bh1;
                    bl[t][e] = bl1;
                }
            }
#pragma unroll
            for (int I = 0; I < NB; ++I)
#pragma unroll
                for (int Jt = 0; Jt < NB; ++Jt) {
                    acc[I][Jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[I], bh[Jt], acc[I][Jt], 0, 0, 0);
                    acc[I][Jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[I], bl[Jt], acc[I][Jt], 0, 0, 0);
                    acc[I][Jt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[I], bh[Jt], acc[I][Jt], 0, 0, 0);
                }
        }
#pragma unroll
        for (int I = 0; I < NB; ++I)
#pragma unroll
            for (int Jt = 0; Jt < NB; ++Jt)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    blk_store(out, k, cI + 16 * I + 4 * kq + r, cJ + 16 * Jt + lr,
                              (double)acc[I][Jt][r]);
    }
    if (tblk == 0) {
        p3 = wave_sum(p3);
        if (lane == 0) out[(size_t)k * k + k + 1] = p3;
    }
}

// LDS of the eigen/assembly kernel (doubles)
struct EigLds {
    double* Z;    // k x lda: C, then Householder vectors + Q, then eigenvectors (columns)
    double* d;    // k  diagonal -> eigenvalues
    double* e;    // k  sub-diagonal
    double* tau;  // k  Householder scalars
    double* rot;  // 2k (c, s) of one QL sweep
    double* B;    // k  Y^T R^-1 d
    double* U;    // k
    double* Wb;   // k  w-bar
    double* sc;   // 8
    double* X;    // nvmax * k  perturbation rows
    double* R;    // nvmax rows of k (k + 1 apart in the large-k vector path)
    double* M;    // nvmax  row means
    double* Cw;   // nvmax  X' w-bar
    double* sw;   // k  sqrt of the obs weights (large-k dual form)
    double* vb;   // 64 Householder vector of a step (n <= 64 reduction)
    double* wb;   // 64 its w = p - K v
};

__host__ __device__ inline int eig_lda(int k) { return k | 1; }

__host__ __device__ inline size_t eig_lds_doubles(int k, int nvmax)
{
    return (size_t)k * eig_lda(k) + 10 * (size_t)k + 8 + 2 * (size_t)nvmax * k + 3 * (size_t)nvmax +
           128;
}

// Z in LDS (zg == nullptr) or, for the large-k kernel, the point's workspace
// record in global memory (zg)
__host__ __device__ inline size_t eig_lds_doubles_big(int k, int nvmax)
{
    return 10 * (size_t)k + 8 + 2 * (size_t)nvmax * k + 3 * (size_t)nvmax + 128;
}

// the split path's tridiagonalisation uses only Z, d, e, tau, vb and wb
__host__ __device__ inline long tri_lds_doubles(int k) { return (long)k * eig_lda(k) + 3L * k + 128; }

__device__ inline EigLds eig_carve_tri(double* s, int k)
{
    EigLds L = {};
    L.Z = s;
    s += (size_t)k * eig_lda(k);
    L.d = s;
    s += k;
    L.e = s;
    s += k;
    L.tau = s;
    s += k;
    L.vb = s;
    s += 64;
    L.wb = s;
    return L;
}

__device__ inline EigLds eig_carve(double* s, int k, int nvmax, double* zg = nullptr)
{
    EigLds L;
    if (zg) {
        L.Z = zg;
    } else {
        L.Z = s;
        s += (size_t)k * eig_lda(k);
    }
    L.d = s;
    s += k;
    L.e = s;
    s += k;
    L.tau = s;
    s += k;
    L.rot = s;
    s += 2 * k;
    L.B = s;
    s += k;
    L.U = s;
    s += k;
    L.Wb = s;
    s += k;
    L.sc = s;
    s += 8;
    L.X = s;
    s += (size_t)nvmax * k;
    L.R = s;
    s += (size_t)nvmax * (k + 1);
    L.M = s;
    s += nvmax;
    L.Cw = s;
    s += nvmax;
    L.sw = s;
    s += k;
    L.vb = s;
    s += 64;
    L.wb = s;
    return L;
}

// Householder reduction of the symmetric n x n matrix in Z to tridiagonal
// T = Q^T A Q (diagonal d, sub-diagonal e), then Q accumulated in place in Z.
// Step j reflects x = A[j+1.., j] onto beta e_1 with H = I - tau v v^T,
// v_0 = 1, beta = -sign(x_0) |x|; the trailing block is updated as
// A -= v w^T + w v^T with p = tau A v, w = p - (tau/2)(v.p) v.  One wave;
// lane r owns rows/columns r and r + 64 (n <= 128).
__device__ void eig_tridiag(EigLds& L, int n, int lda, int lane)
{
    double* A = L.Z;
    for (int j = 0; j + 2 < n; ++j) {
        const int o = j + 1, m = n - o;
        double x0 = 0.0, x1 = 0.0;  // x_r for r = lane, lane + 64
        if (lane < m) x0 = A[(o + lane) * lda + j];
        if (lane + 64 < m) x1 = A[(o + lane + 64) * lda + j];
        const double sigma = wave_sum((lane > 0 ? x0 * x0 : 0.0) + x1 * x1);
        const double alpha = lane_bcast(x0, 0);
        double tau = 0.0, beta = alpha, scal = 0.0;
        if (sigma > 0.0) {
            const double nrm = sqrt(fma(alpha, alpha, sigma));
            beta = alpha <= 0.0 ? nrm : -nrm;
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        double v0 = lane == 0 ? 1.0 : x0 * scal, v1 = x1 * scal;
        if (tau == 0.0) {
            v0 = lane == 0 ? 1.0 : 0.0;
            v1 = 0.0;
        }
        if (lane == 0) {
            L.d[j] = A[j * lda + j];
            L.e[j] = beta;
            L.tau[j] = tau;
        }
        if (lane < m) A[(o + lane) * lda + j] = v0;
        if (lane + 64 < m) A[(o + lane + 64) * lda + j] = v1;
        __syncthreads();
        if (tau == 0.0) continue;
        // p = tau A22 v (A22 symmetric: lane r reads column r, consecutive over lanes)
        double p0 = 0.0, p1 = 0.0;
#pragma unroll 4
        for (int c = 0; c < m; ++c) {
            const double vc = lane_bcast(c < 64 ? v0 : v1, c & 63);
            const double* row = A + (o + c) * lda + o;
            if (lane < m) p0 = fma(row[lane], vc, p0);
            if (lane + 64 < m) p1 = fma(row[lane + 64], vc, p1);
        }
        p0 *= tau;
        p1 *= tau;
        const double K = 0.5 * tau * wave_sum(p0 * v0 + p1 * v1);
        const double w0 = fma(-K, v0, p0), w1 = fma(-K, v1, p1);
        // A22 -= v w^T + w v^T (lane r updates column r).  Both products are
        // rounded and then added, so elements (c, r) and (r, c) get the same
        // value and A22 stays exactly symmetric.
#pragma unroll 4
        for (int c = 0; c < m; ++c) {
            const double vc = lane_bcast(c < 64 ? v0 : v1, c & 63);
            const double wc = lane_bcast(c < 64 ? w0 : w1, c & 63);
            double* row = A + (o + c) * lda + o;
            if (lane < m)
                row[lane] = rn_sub(row[lane], rn_add(rn_mul(vc, w0), rn_mul(wc, v0)));
            if (lane + 64 < m)
                row[lane + 64] =
                    rn_sub(row[lane + 64], rn_add(rn_mul(vc, w1), rn_mul(wc, v1)));
        }
        __syncthreads();
    }
    if (lane == 0) {
        if (n >= 2) {
            L.d[n - 2] = A[(n - 2) * lda + n - 2];
            L.e[n - 2] = A[(n - 1) * lda + n - 2];
        }
        L.d[n - 1] = A[(n - 1) * lda + n - 1];
        L.e[n - 1] = 0.0;
    }
    __syncthreads();
    // Q = H_0 H_1 ... H_{n-3}, accumulated from the right end in place: before
    // step j, row and column j+1 become unit vectors (the reflector stored
    // there has been used), then Q[j+1.., j+1..] -= tau v (v^T Q[j+1.., j+1..]).
    const int t0 = n >= 2 ? n - 2 : 0;
    for (int r = lane; r < n - t0; r += 64)
        for (int c = 0; c < n - t0; ++c) A[(t0 + c) * lda + t0 + r] = (r == c) ? 1.0 : 0.0;
    __syncthreads();
    for (int j = n - 3; j >= 0; --j) {
        const int o = j + 1, m = n - o;
        // unit row/column o
        for (int r = lane; r < m; r += 64) {
            A[o * lda + o + r] = r == 0 ? 1.0 : 0.0;
            A[(o + r) * lda + o] = r == 0 ? 1.0 : 0.0;
        }
        __syncthreads();
        const double tau = L.tau[j];
        if (tau == 0.0) continue;
        double v0 = 0.0, v1 = 0.0;
        if (lane < m) v0 = A[(o + lane) * lda + j];
        if (lane + 64 < m) v1 = A[(o + lane + 64) * lda + j];
        double y0 = 0.0, y1 = 0.0;
#pragma unroll 4
        for (int c = 0; c < m; ++c) {
            const double vc = lane_bcast(c < 64 ? v0 : v1, c & 63);
            const double* row = A + (o + c) * lda + o;
            if (lane < m) y0 = fma(vc, row[lane], y0);
            if (lane + 64 < m) y1 = fma(vc, row[lane + 64], y1);
        }
        y0 *= tau;
        y1 *= tau;
#pragma unroll 4
        for (int c = 0; c < m; ++c) {
            const double vc = lane_bcast(c < 64 ? v0 : v1, c & 63);
            double* row = A + (o + c) * lda + o;
            if (lane < m) row[lane] = fma(-vc, y0, row[lane]);
            if (lane + 64 < m) row[lane + 64] = fma(-vc, y1, row[lane + 64]);
        }
        __syncthreads();
    }
    // row/column 0
    for (int r = lane; r < n; r += 64) {
        A[r] = r == 0 ? 1.0 : 0.0;
        A[r * lda] = r == 0 ? 1.0 : 0.0;
    }
    __syncthreads();
}

// The reduction of eig_tridiag for n <= 64, lane r owning row/column r of the
// trailing block, with the step's v and w in LDS buffers: every inner loop is
// a column read plus broadcast reads (no cross-lane shuffles), unrolled with
// independent loads, and a step's work is its trailing block's.  The matrix-
// vector product sums in four interleaved partial sums.
__device__ void eig_tridiag64(EigLds& L, int n, int lda, int lane)
{
    double* A = L.Z;
    double* vb = L.vb;
    double* wb = L.wb;
    for (int j = 0; j + 2 < n; ++j) {
        const int o = j + 1, m = n - o;
        const double x0 = lane < m ? A[(o + lane) * lda + j] : 0.0;
        const double sigma = wave_sum(lane > 0 ? x0 * x0 : 0.0);
        const double alpha = lane_bcast(x0, 0);
        double tau = 0.0, beta = alpha, scal = 0.0;
        if (sigma > 0.0) {
            const double nrm = sqrt(fma(alpha, alpha, sigma));
            beta = alpha <= 0.0 ? nrm : -nrm;
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        double v0 = lane == 0 ? 1.0 : x0 * scal;
        if (tau == 0.0) v0 = lane == 0 ? 1.0 : 0.0;
        if (lane == 0) {
            L.d[j] = A[j * lda + j];
            L.e[j] = beta;
            L.tau[j] = tau;
        }
        if (lane < m) {
            A[(o + lane) * lda + j] = v0;
            vb[lane] = v0;
        }
        __syncthreads();
        if (tau == 0.0) continue;
        // p = tau A22 v: lane r reads column r (= row r) of the trailing block
        double p = 0.0;
        if (lane < m) {
            const double* col = A + o * lda + o + lane;
            double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
            int c = 0;
            for (; c + 4 <= m; c += 4) {
                p0 = fma(col[c * lda], vb[c], p0);
                p1 = fma(col[(c + 1) * lda], vb[c + 1], p1);
                p2 = fma(col[(c + 2) * lda], vb[c + 2], p2);
                p3 = fma(col[(c + 3) * lda], vb[c + 3], p3);
            }
            for (; c < m; ++c) p0 = fma(col[c * lda], vb[c], p0);
            p = tau * ((p0 + p1) + (p2 + p3));
        }
        const double K = 0.5 * tau * wave_sum(lane < m ? v0 * p : 0.0);
        const double w0 = fma(-K, v0, p);
        if (lane < m) wb[lane] = w0;
        __syncthreads();
        // A22 -= v w^T + w v^T (lane r updates column r; both products are
        // rounded, then added, so (c, r) and (r, c) stay equal).  A group's
        // loads are all issued before its stores: v, w and the column share
        // LDS, so element by element each load would wait for the store before
        if (lane < m) {
            double* col = A + o * lda + o + lane;
            int c = 0;
            for (; c + 8 <= m; c += 8) {
                double cv[8], vv[8], ww[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    cv[u] = col[(c + u) * lda];
                    vv[u] = vb[c + u];
                    ww[u] = wb[c + u];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    col[(c + u) * lda] = rn_sub(cv[u], rn_add(rn_mul(vv[u], w0), rn_mul(ww[u], v0)));
            }
            if (c + 4 <= m) {
                double cv[4], vv[4], ww[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    cv[u] = col[(c + u) * lda];
                    vv[u] = vb[c + u];
                    ww[u] = wb[c + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    col[(c + u) * lda] = rn_sub(cv[u], rn_add(rn_mul(vv[u], w0), rn_mul(ww[u], v0)));
                c += 4;
            }
            for (; c < m; ++c)
                col[c * lda] = rn_sub(col[c * lda], rn_add(rn_mul(vb[c], w0), rn_mul(wb[c], v0)));
        }
        __syncthreads();
    }
    if (lane == 0) {
        if (n >= 2) {
            L.d[n - 2] = A[(n - 2) * lda + n - 2];
            L.e[n - 2] = A[(n - 1) * lda + n - 2];
        }
        L.d[n - 1] = A[(n - 1) * lda + n - 1];
        L.e[n - 1] = 0.0;
    }
    __syncthreads();
    // Q = H_0 H_1 ... H_{n-3}, accumulated from the right end in place (as
    // eig_tridiag): before step j, row and column j+1 become unit vectors, then
    // Q[j+1.., j+1..] -= tau v (v^T Q[j+1.., j+1..]).
    const int t0 = n >= 2 ? n - 2 : 0;
    if (lane < n - t0)
        for (int c = 0; c < n - t0; ++c) A[(t0 + c) * lda + t0 + lane] = (lane == c) ? 1.0 : 0.0;
    __syncthreads();
    for (int j = n - 3; j >= 0; --j) {
        const int o = j + 1, m = n - o;
        if (lane < m) {
            A[o * lda + o + lane] = lane == 0 ? 1.0 : 0.0;
            A[(o + lane) * lda + o] = lane == 0 ? 1.0 : 0.0;
        }
        const double tau = L.tau[j];
        if (tau != 0.0 && lane < m) vb[lane] = A[(o + lane) * lda + j];
        __syncthreads();
        if (tau == 0.0) continue;
        if (lane < m) {
            double* col = A + o * lda + o + lane;
            double y0 = 0.0, y1 = 0.0, y2 = 0.0, y3 = 0.0;
            int c = 0;
            for (; c + 4 <= m; c += 4) {
                y0 = fma(vb[c], col[c * lda], y0);
                y1 = fma(vb[c + 1], col[(c + 1) * lda], y1);
                y2 = fma(vb[c + 2], col[(c + 2) * lda], y2);
                y3 = fma(vb[c + 3], col[(c + 3) * lda], y3);
            }
            for (; c < m; ++c) y0 = fma(vb[c], col[c * lda], y0);
            const double y = tau * ((y0 + y1) + (y2 + y3));
            c = 0;
            // loads of a group before its stores (v and Q share LDS)
            for (; c + 8 <= m; c += 8) {
                double cv[8], vv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    cv[u] = col[(c + u) * lda];
                    vv[u] = vb[c + u];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) col[(c + u) * lda] = fma(-vv[u], y, cv[u]);
            }
            if (c + 4 <= m) {
                double cv[4], vv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    cv[u] = col[(c + u) * lda];
                    vv[u] = vb[c + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) col[(c + u) * lda] = fma(-vv[u], y, cv[u]);
                c += 4;
            }
            for (; c < m; ++c) col[c * lda] = fma(-vb[c], y, col[c * lda]);
        }
        __syncthreads();
    }
    if (lane < n) {
        A[lane] = lane == 0 ? 1.0 : 0.0;
        A[lane * lda] = lane == 0 ? 1.0 : 0.0;
    }
    __syncthreads();
}

// The same reduction for n <= 64 * NR with the matrix in global memory
// (row-major, leading dimension lda, exactly symmetric).  Lane r owns
// columns r, r + 64, ... of the trailing block.  The reflector of step j is
// read from and kept in row j (equal to column j by symmetry), so every
// access is a contiguous row segment.  Q is accumulated as above and then
// transposed in place: on return Z[r + c * lda] = Q[r][c], the layout the
// lane-per-row QL sweeps read contiguously.
constexpr int kRowBatch = 8;  // rows in flight per lane in the global-memory reduction

template <int NR, bool FORMQ = true>
__device__ void big_tridiag(double* A, int n, int lda, double* d, double* e, double* tauv, int lane)
{
    for (int j = 0; j + 2 < n; ++j) {
        const int o = j + 1, m = n - o;
        double* rowj = A + (size_t)j * lda + o;
        double x[NR];
        double s2 = 0.0;
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            const int r = lane + 64 * t;
            x[t] = r < m ? rowj[r] : 0.0;
            if (r > 0) s2 = fma(x[t], x[t], s2);
        }
        const double sigma = wave_sum(s2);
        const double alpha = lane_bcast(x[0], 0);
        double tau = 0.0, beta = alpha, scal = 0.0;
        if (sigma > 0.0) {
            const double nrm = sqrt(fma(alpha, alpha, sigma));
            beta = alpha <= 0.0 ? nrm : -nrm;
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        double v[NR];
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            const int r = lane + 64 * t;
            v[t] = r == 0 ? 1.0 : (tau == 0.0 ? 0.0 : x[t] * scal);
        }
        if (lane == 0) {
            d[j] = A[(size_t)j * lda + j];
            e[j] = beta;
            tauv[j] = tau;
        }
#pragma unroll
        for (int t = 0; t < NR; ++t)
            if (lane + 64 * t < m) rowj[lane + 64 * t] = v[t];
        __syncthreads();
        if (tau == 0.0) continue;
        double pp[NR];
#pragma unroll
        for (int t = 0; t < NR; ++t) pp[t] = 0.0;
        // rows are loaded kRowBatch at a time so that many loads are in flight
#pragma unroll
        for (int cb = 0; cb < NR; ++cb) {
            const int cn = m - 64 * cb < 64 ? m - 64 * cb : 64;
            if (cn <= 0) break;
            for (int c0 = 0; c0 < cn; c0 += kRowBatch) {
                double x[kRowBatch][NR];
#pragma unroll
                for (int u = 0; u < kRowBatch; ++u) {
                    const double* row = A + (size_t)(o + 64 * cb + c0 + u) * lda + o;
#pragma unroll
                    for (int t = 0; t < NR; ++t)
                        x[u][t] = (c0 + u < cn && lane + 64 * t < m) ? row[lane + 64 * t] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < kRowBatch; ++u) {
                    const double vc = lane_bcast(v[cb], (c0 + u) & 63);
#pragma unroll
                    for (int t = 0; t < NR; ++t) pp[t] = fma(x[u][t], vc, pp[t]);
                }
            }
        }
        double pv = 0.0;
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            pp[t] *= tau;
            pv = fma(pp[t], v[t], pv);
        }
        const double K = 0.5 * tau * wave_sum(pv);
        double w[NR];
#pragma unroll
        for (int t = 0; t < NR; ++t) w[t] = fma(-K, v[t], pp[t]);
#pragma unroll
        for (int cb = 0; cb < NR; ++cb) {
            const int cn = m - 64 * cb < 64 ? m - 64 * cb : 64;
            if (cn <= 0) break;
            for (int c0 = 0; c0 < cn; c0 += kRowBatch) {
                double x[kRowBatch][NR];
#pragma unroll
                for (int u = 0; u < kRowBatch; ++u) {
                    const double* row = A + (size_t)(o + 64 * cb + c0 + u) * lda + o;
#pragma unroll
                    for (int t = 0; t < NR; ++t)
                        x[u][t] = (c0 + u < cn && lane + 64 * t < m) ? row[lane + 64 * t] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < kRowBatch; ++u) {
                    const double vc = lane_bcast(v[cb], (c0 + u) & 63);
                    const double wc = lane_bcast(w[cb], (c0 + u) & 63);
                    double* row = A + (size_t)(o + 64 * cb + c0 + u) * lda + o;
#pragma unroll
                    for (int t = 0; t < NR; ++t)
                        if (c0 + u < cn && lane + 64 * t < m)
                            row[lane + 64 * t] = rn_sub(
                                x[u][t], rn_add(rn_mul(vc, w[t]), rn_mul(wc, v[t])));
                }
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        d[n - 2] = A[(size_t)(n - 2) * lda + n - 2];
        e[n - 2] = A[(size_t)(n - 1) * lda + n - 2];
        d[n - 1] = A[(size_t)(n - 1) * lda + n - 1];
        e[n - 1] = 0.0;
    }
    __syncthreads();
    if (!FORMQ) return;
    // Q = H_0 ... H_{n-3} from the right end, as in eig_tridiag; the reflector
    // of step j is in row j, columns j+1..
    for (int r = lane; r < 4; r += 64) {
        const int i = r >> 1, c = r & 1;
        A[(size_t)(n - 2 + i) * lda + n - 2 + c] = i == c ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int j = n - 3; j >= 0; --j) {
        const int o = j + 1, m = n - o;
        for (int r = lane; r < m; r += 64) {
            A[(size_t)o * lda + o + r] = r == 0 ? 1.0 : 0.0;
            A[(size_t)(o + r) * lda + o] = r == 0 ? 1.0 : 0.0;
        }
        __syncthreads();
        const double tau = tauv[j];
        if (tau == 0.0) continue;
        double v[NR], y[NR];
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            const int r = lane + 64 * t;
            v[t] = r < m ? A[(size_t)j * lda + o + r] : 0.0;
            y[t] = 0.0;
        }
#pragma unroll
        for (int cb = 0; cb < NR; ++cb) {
            const int cn = m - 64 * cb < 64 ? m - 64 * cb : 64;
            if (cn <= 0) break;
#pragma unroll 16
            for (int cc = 0; cc < cn; ++cc) {
                const double vc = lane_bcast(v[cb], cc);
                const double* row = A + (size_t)(o + 64 * cb + cc) * lda + o;
#pragma unroll
                for (int t = 0; t < NR; ++t)
                    if (lane + 64 * t < m) y[t] = fma(vc, row[lane + 64 * t], y[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < NR; ++t) y[t] *= tau;
#pragma unroll
        for (int cb = 0; cb < NR; ++cb) {
            const int cn = m - 64 * cb < 64 ? m - 64 * cb : 64;
            if (cn <= 0) break;
#pragma unroll 16
            for (int cc = 0; cc < cn; ++cc) {
                const double vc = lane_bcast(v[cb], cc);
                double* row = A + (size_t)(o + 64 * cb + cc) * lda + o;
#pragma unroll
                for (int t = 0; t < NR; ++t)
                    if (lane + 64 * t < m) row[lane + 64 * t] = fma(-vc, y[t], row[lane + 64 * t]);
            }
        }
        __syncthreads();
    }
    for (int r = lane; r < n; r += 64) {
        A[r] = r == 0 ? 1.0 : 0.0;
        A[(size_t)r * lda] = r == 0 ? 1.0 : 0.0;
    }
    __syncthreads();
    // in-place transpose: row i, columns i+1.. swap with column i
    for (int i = 0; i + 1 < n; ++i) {
        for (int c = i + 1 + lane; c < n; c += 64) {
            const double a = A[(size_t)i * lda + c], bb = A[(size_t)c * lda + i];
            A[(size_t)i * lda + c] = bb;
            A[(size_t)c * lda + i] = a;
        }
    }
    __syncthreads();
}

// The same reduction for the vector paths, with a third of the memory
// traffic: only the upper triangle (row c, columns >= c) is kept, and step
// s's rank-2 update is applied during step s+1's single pass over the
// trailing block, which also forms step s+1's product B v (p_r gets
// A[c][r] v_c from rows c < r and the row-r sums, reduced across lanes).
// vP, wP (pending update) and vN (current reflector) are LDS vectors of n
// indexed absolutely.  On exit d, e, tau and the reflectors (row s,
// columns s+1..) are as big_tridiag leaves them, the elementwise update
// being rounded identically.
template <int NR>
__device__ void big_tridiag_sym(double* A, int n, int lda, double* d, double* e, double* tauv,
                                double* vP, double* wP, double* vN, int lane)
{
    for (int r = lane; r < n; r += 64) {
        vP[r] = 0.0;
        wP[r] = 0.0;
    }
    __syncthreads();
    for (int s = 0; s + 2 < n; ++s) {
        const int o = s + 1, m = n - o;
        // row s with the pending update: diagonal and x = A[s][o..]
        const double vps = vP[s], wps = wP[s];
        double* rows = A + (size_t)s * lda;
        double x[NR];
        double s2 = 0.0;
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            const int i = lane + 64 * t, r = o + i;
            x[t] = 0.0;
            if (i < m)
                x[t] = rn_sub(rows[r], rn_add(rn_mul(vps, wP[r]), rn_mul(wps, vP[r])));
            if (i > 0) s2 = fma(x[t], x[t], s2);
        }
        const double dss = rn_sub(rows[s], rn_add(rn_mul(vps, wps), rn_mul(wps, vps)));
        const double sigma = wave_sum(s2);
        const double alpha = lane_bcast(x[0], 0);
        double tau = 0.0, beta = alpha, scal = 0.0;
        if (sigma > 0.0) {
            const double nrm = sqrt(fma(alpha, alpha, sigma));
            beta = alpha <= 0.0 ? nrm : -nrm;
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        double v[NR];
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            const int i = lane + 64 * t;
            v[t] = i == 0 ? 1.0 : (tau == 0.0 ? 0.0 : x[t] * scal);
            if (i < m) {
                rows[o + i] = v[t];
                vN[o + i] = v[t];
            }
        }
        if (lane == 0) {
            d[s] = dss;
            e[s] = beta;
            tauv[s] = tau;
        }
        __syncthreads();
        // one pass over rows c = o.. (columns r >= c): pending update, store,
        // and p = B v of this step
        double wpr[NR], vpr[NR], vnr[NR], pp[NR];
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            const int i = lane + 64 * t, r = o + i;
            wpr[t] = i < m ? wP[r] : 0.0;
            vpr[t] = i < m ? vP[r] : 0.0;
            vnr[t] = i < m ? v[t] : 0.0;
            pp[t] = 0.0;
        }
#pragma unroll
        for (int cb = 0; cb < NR; ++cb) {
            const int cn = m - 64 * cb < 64 ? m - 64 * cb : 64;
            if (cn <= 0) break;
            for (int c0 = 0; c0 < cn; c0 += kRowBatch) {
                double a[kRowBatch][NR];
#pragma unroll
                for (int u = 0; u < kRowBatch; ++u) {
                    const int ic = 64 * cb + c0 + u;  // row index relative to o
                    const double* row = A + (size_t)(o + ic) * lda + o;
#pragma unroll
                    for (int t = 0; t < NR; ++t) {
                        const int i = lane + 64 * t;
                        a[u][t] = (c0 + u < cn && t >= cb && i >= ic && i < m) ? row[i] : 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < kRowBatch; ++u) {
                    const int ic = 64 * cb + c0 + u;
                    const bool rv = c0 + u < cn;
                    const int c = o + (rv ? ic : 0);
                    const double vpc = vP[c], wpc = wP[c], vnc = vN[c];
                    double* row = A + (size_t)(o + ic) * lda + o;
                    double rs = 0.0;
#pragma unroll
                    for (int t = 0; t < NR; ++t) {
                        const int i = lane + 64 * t;
                        if (rv && t >= cb && i >= ic && i < m) {
                            const double an = rn_sub(
                                a[u][t], rn_add(rn_mul(vpc, wpr[t]), rn_mul(wpc, vpr[t])));
                            row[i] = an;
                            if (i > ic) {
                                pp[t] = fma(an, vnc, pp[t]);
                                rs = fma(an, vnr[t], rs);
                            } else {
                                rs = fma(an, vnc, rs);
                            }
                        }
                    }
                    rs = wave_sum(rs);
                    if (rv && lane == ((c0 + u) & 63)) pp[cb] += rs;
                }
            }
        }
        __syncthreads();  // every lane is done reading this step's vP, wP, vN
        double pv = 0.0;
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            pp[t] *= tau;
            pv = fma(pp[t], v[t], pv);
        }
        const double K = 0.5 * tau * wave_sum(pv);
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            const int i = lane + 64 * t;
            if (i < m) {
                vP[o + i] = v[t];
                wP[o + i] = fma(-K, v[t], pp[t]);
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        const int a0 = n - 2, a1 = n - 1;
        d[a0] = rn_sub(A[(size_t)a0 * lda + a0],
                          rn_add(rn_mul(vP[a0], wP[a0]), rn_mul(wP[a0], vP[a0])));
        e[a0] = rn_sub(A[(size_t)a0 * lda + a1],
                          rn_add(rn_mul(vP[a0], wP[a1]), rn_mul(wP[a0], vP[a1])));
        d[a1] = rn_sub(A[(size_t)a1 * lda + a1],
                          rn_add(rn_mul(vP[a1], wP[a1]), rn_mul(wP[a1], vP[a1])));
        e[a1] = 0.0;
    }
    __syncthreads();
}

// sweeps allowed per eigenvalue before the solve gives up, as tql2's 30
// (common/netlib.f:718-887; the reference then STOPs, common_mtx.f90:61-64)
constexpr int kQlMaxIter = 30;

// ql_apply for Z in global memory with Z[r + i * lda] = Z_ri (rows of a
// column contiguous), NR rows per lane
template <int NR>
__device__ __forceinline__ void ql_apply_cm(EigLds& L, int n, int lda, int lane, int ilo, int m)
{
    double z1[NR];
#pragma unroll
    for (int t = 0; t < NR; ++t) {
        const int r = lane + 64 * t;
        z1[t] = r < n ? L.Z[r + (size_t)m * lda] : 0.0;
    }
    for (int i = m - 1; i >= ilo; --i) {
        const double c = L.rot[2 * i], s = L.rot[2 * i + 1];
        const double* zi = L.Z + (size_t)i * lda;
        double* zi1 = L.Z + (size_t)(i + 1) * lda;
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            const int r = lane + 64 * t;
            if (r < n) {
                const double z0 = zi[r];
                zi1[r] = fma(s, z0, c * z1[t]);
                z1[t] = fma(c, z0, -s * z1[t]);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NR; ++t) {
        const int r = lane + 64 * t;
        if (r < n) L.Z[r + (size_t)ilo * lda] = z1[t];
    }
}

// Rotations (c_i, s_i), i = m-1 down to ilo, of one QL sweep applied to the
// columns (i, i+1) of Z, row by row (lane = row): the new z_{i+1} is stored,
// the new z_i carried to the next step.
__device__ __forceinline__ void ql_apply(EigLds& L, int n, int lda, int lane, int ilo, int m)
{
    for (int r = lane; r < n; r += 64) {
        double* z = L.Z + r * lda;
        double z1 = z[m];
#pragma unroll 4
        for (int i = m - 1; i >= ilo; --i) {
            const double c = L.rot[2 * i], s = L.rot[2 * i + 1];
            const double z0 = z[i];
            z[i + 1] = fma(s, z0, c * z1);
            z1 = fma(c, z0, -s * z1);
        }
        z[ilo] = z1;
    }
}

// Eigenvalues (into d) and eigenvectors (columns of Z, which holds Q on
// entry) of the symmetric tridiagonal (d, e) by implicit QL with shifts
// (Bowdler, Martin, Reinsch and Wilkinson 1968, the method of EISPACK tql2).
// A sub-diagonal element is negligible when adding it to the running norm
// estimate tst = max(|d_l| + |e_l|) does not change it.  Every lane runs
// the scalar recurrence; the rotations of a sweep are kept in L.rot and
// then applied to Z by all lanes.  Returns nonzero if an eigenvalue needed
// more than kQlMaxIter sweeps.  CM > 0: Z in global memory, column-major
// (ql_apply_cm with CM rows per lane).
template <int CM = 0>
__device__ int eig_ql(EigLds& L, int n, int lda, int lane)
{
    double* d = L.d;
    double* e = L.e;
    double tst = 0.0;
    for (int l = 0; l < n; ++l) {
        tst = fmax(tst, fabs(d[l]) + fabs(e[l]));
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; ++m)
                if (tst + fabs(e[m]) == tst) break;
            if (m == l) break;
            if (iter++ == kQlMaxIter) return 1;
            const double dl = d[l], el = e[l];
            double g = (d[l + 1] - dl) / (2.0 * el);
            double r = fabs(g) > 1.0e100 ? fabs(g) : sqrt(fma(g, g, 1.0));
            g = d[m] - dl + el / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            bool split = false;
            double ei = e[i], di1 = d[i + 1];
            for (; i >= l; --i) {
                const double di = d[i];
                const double eprev = i > l ? e[i - 1] : 0.0;  // next step's e, loaded early
                const double f = s * ei, bb = c * ei;
                r = sqrt(fma(f, f, g * g));
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] = di1 - p;
                    e[m] = 0.0;
                    split = true;
                    break;
                }
                const double ri = rcp_nr(r);
                s = f * ri;
                c = g * ri;
                g = di1 - p;
                r = fma(di - g, s, 2.0 * c * bb);
                p = s * r;
                d[i + 1] = g + p;
                g = fma(c, r, -bb);
                if (lane == 0) {
                    L.rot[2 * i] = c;
                    L.rot[2 * i + 1] = s;
                }
                ei = eprev;
                di1 = di;
            }
            __syncthreads();
            if constexpr (CM > 0)
                ql_apply_cm<CM>(L, n, lda, lane, split ? i + 1 : l, m);
            else
                ql_apply(L, n, lda, lane, split ? i + 1 : l, m);
            if (!split) {
                d[l] = d[l] - p;
                e[l] = g;
                e[m] = 0.0;
            }
            __syncthreads();
        }
    }
    return 0;
}

// Householder + implicit-QL eigendecomposition for n <= KR <= 64: the
// tridiagonalisation in LDS (eig_tridiag64), then QL with Q in registers: lane
// r holds row r of Q, later of Z (q[t] = Z[r][t]); d_i and e_i live in lane
// i and are read by the others with readlane; a sweep runs the scalar
// recurrence in every lane and rotates columns (i, i+1) of the lane's row as
// it goes (loops over matrix indices run over the static index t with
// uniform branches, so every register index is a compile-time constant).
// Reads C (+ the diagonal shift) from Z in LDS, writes the eigenvectors back
// to Z (columns) and the eigenvalues to d.  Returns nonzero if QL did not
// converge.
// Implicit QL with shifts on the tridiagonal (d_i, e_i in lane i), rotating
// the rows q[] of Q as each rotation is formed.  Returns nonzero if an
// eigenvalue took more than kQlMaxIter sweeps.
template <int KR>
__device__ __forceinline__ int ql_reg(double (&q)[KR], int n, int lane, double& dreg,
                                      double& ereg)
{
    double tst = 0.0;
    int bad = 0;
    for (int l = 0; l < n && !bad; ++l) {
        tst = fmax(tst, fabs(lane_bcast(dreg, l)) + fabs(lane_bcast(ereg, l)));
        for (int iter = 0;; ++iter) {
            const bool negl = lane >= l && lane < n - 1 && (tst + fabs(ereg) == tst);
            const unsigned long long mask =
                (__ballot(negl) | (1ull << (n - 1))) & (~0ull << l);
            const int m = __builtin_ctzll(mask);
            if (m == l) break;
            if (iter == kQlMaxIter) {
                bad = 1;
                break;
            }
            const double dl = lane_bcast(dreg, l), el = lane_bcast(ereg, l);
            double g = (lane_bcast(dreg, l + 1) - dl) / (2.0 * el);
            double r = fabs(g) > 1.0e100 ? fabs(g) : sqrt(fma(g, g, 1.0));
            g = lane_bcast(dreg, m) - dl + el / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0, di1 = lane_bcast(dreg, m);
            bool split = false;
#pragma unroll
            for (int i = KR - 2; i >= 0; --i) {
                if (i >= l && i < m && !split) {
                    const double ei = lane_bcast(ereg, i), di = lane_bcast(dreg, i);
                    const double f = s * ei, bb = c * ei;
                    const double x = fma(f, f, g * g);
                    if (x == 0.0) {
                        if (lane == i + 1) dreg = di1 - p;
                        if (lane == m) ereg = 0.0;
                        split = true;
                    } else {
                        const double y = rsq_nr(x);
                        r = x * y;  // hypot(f, g)
                        s = f * y;
                        c = g * y;
                        if (lane == i + 1) ereg = r;
                        g = di1 - p;
                        r = fma(di - g, s, 2.0 * c * bb);
                        p = s * r;
                        if (lane == i + 1) dreg = g + p;
                        g = fma(c, r, -bb);
                        di1 = di;
                        const double h = q[i + 1];
                        q[i + 1] = fma(s, q[i], c * h);
                        q[i] = fma(c, q[i], -s * h);
                    }
                }
            }
            if (!split) {
                if (lane == l) {
                    dreg -= p;
                    ereg = g;
                }
                if (lane == m) ereg = 0.0;
            }
        }
    }
    return bad;
}

template <int KR>
__device__ int eig_reg(EigLds& L, int n, int lda, int lane)
{
    // Householder in LDS (each step's work is the active block's), then Q's
    // row r into lane r's registers for the QL sweeps
    eig_tridiag64(L, n, lda, lane);
    double q[KR];
#pragma unroll
    for (int t = 0; t < KR; ++t) q[t] = (t < n && lane < n) ? L.Z[lane * lda + t] : 0.0;
    double dreg = lane < n ? L.d[lane] : 0.0, ereg = lane < n ? L.e[lane] : 0.0;
    __syncthreads();
    const int bad = ql_reg<KR>(q, n, lane, dreg, ereg);
    // eigenvalues and eigenvectors (columns of Z) to LDS
    if (lane < n) {
        L.d[lane] = dreg;
#pragma unroll
        for (int t = 0; t < KR; ++t)
            if (t < n) L.Z[lane * lda + t] = q[t];
    }
    __syncthreads();
    return bad;
}

// ---------------------------------------------------------------------------
// Large-k path without the trans/pao outputs: the eigenvectors V = Q S are
// never formed.  The analysis needs V only through V^T y and V c for a few
// vectors (b, the point's perturbation rows, w-bar), so those vectors are
// carried through Q^T (the reflectors) and through each QL sweep's rotations
// as the sweeps are computed (S^T), and the rotations are kept in a buffer
// after the point's workspace record to be replayed in reverse (S) before Q is
// applied.  Traffic per point drops from O(k^3) matrix sweeps to the
// reduction itself plus O(k^2) per vector.
// ---------------------------------------------------------------------------

// rotation buffer: (c, s) pairs of up to 2k^2 + 4k rotations plus one
// (ilo, m) trailer per sweep
__host__ __device__ inline long big_rot_cap(int kmax) { return 2 * (2L * kmax * kmax + 4L * kmax); }
__host__ __device__ inline long big_stride(int kmax) { return gram_stride(kmax) + big_rot_cap(kmax); }

// the vectors carried along: nrow rows of R (k apart), then v1, then v2
struct VecSet {
    double* R;
    int nrow, k;
    double* v1;
    double* v2;
    __device__ double* at(int q) const
    {
        return q < nrow ? R + (size_t)q * k : (q == nrow ? v1 : v2);
    }
};

// y <- Q^T y (transpose: H_0 first) or y <- Q y (H_{n-3} first) for nvec
// vectors, H_j = I - tau_j v_j v_j^T with v_j in row j, columns j+1.., of A.
// Lane r owns elements r, r + 64, ... of every vector.
template <int NR>
__device__ void big_refl_vecs(const double* A, int n, int lda, const double* tauv, const VecSet& V,
                              int nvec, bool transpose, int lane)
{
    for (int s = 0; s + 2 < n; ++s) {
        const int j = transpose ? s : n - 3 - s;
        const double tau = tauv[j];
        if (tau == 0.0) continue;
        const int o = j + 1, m = n - o;
        double v[NR];
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            const int r = lane + 64 * t;
            v[t] = r < m ? A[(size_t)j * lda + o + r] : 0.0;
        }
        for (int q = 0; q < nvec; ++q) {
            double* y = V.at(q) + o;
            double sum = 0.0;
#pragma unroll
            for (int t = 0; t < NR; ++t)
                if (lane + 64 * t < m) sum = fma(v[t], y[lane + 64 * t], sum);
            sum = tau * wave_sum(sum);
#pragma unroll
            for (int t = 0; t < NR; ++t)
                if (lane + 64 * t < m) y[lane + 64 * t] = fma(-sum, v[t], y[lane + 64 * t]);
        }
        // element o + r belongs to lane r, so ownership shifts by one lane per
        // reflector: order this step's LDS stores before the next step's loads
        __syncthreads();
    }
}

// eig_ql on (d, e) without eigenvectors: each sweep's rotations are applied
// to the nvec vectors (lane q owns vector q; the same update as a row of Z,
// so u = Z^T y is maintained) and appended to rbuf.  Returns nonzero if an
// eigenvalue needed more than kQlMaxIter sweeps; ovf is set if rbuf filled.
__device__ int eig_ql_vec(EigLds& L, int n, int lane, const VecSet& V, int nvec, double* rbuf,
                          long rcap, long& rpos, bool& ovf)
{
    double* d = L.d;
    double* e = L.e;
    double tst = 0.0;
    for (int l = 0; l < n; ++l) {
        tst = fmax(tst, fabs(d[l]) + fabs(e[l]));
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; ++m)
                if (tst + fabs(e[m]) == tst) break;
            if (m == l) break;
            if (iter++ == kQlMaxIter) return 1;
            const double dl = d[l], el = e[l];
            double g = (d[l + 1] - dl) / (2.0 * el);
            double r = fabs(g) > 1.0e100 ? fabs(g) : sqrt(fma(g, g, 1.0));
            g = d[m] - dl + el / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            bool split = false;
            double ei = e[i], di1 = d[i + 1];
            for (; i >= l; --i) {
                const double di = d[i];
                const double eprev = i > l ? e[i - 1] : 0.0;
                const double f = s * ei, bb = c * ei;
                r = sqrt(fma(f, f, g * g));
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] = di1 - p;
                    e[m] = 0.0;
                    split = true;
                    break;
                }
                const double ri = rcp_nr(r);
                s = f * ri;
                c = g * ri;
                g = di1 - p;
                r = fma(di - g, s, 2.0 * c * bb);
                p = s * r;
                d[i + 1] = g + p;
                g = fma(c, r, -bb);
                if (lane == 0) {
                    L.rot[2 * i] = c;
                    L.rot[2 * i + 1] = s;
                }
                ei = eprev;
                di1 = di;
            }
            __syncthreads();
            const int ilo = split ? i + 1 : l;
            const int cnt = m - ilo;
            if (!ovf && rpos + 2L * cnt + 2 <= rcap) {
                for (int q = lane; q < 2 * cnt; q += 64)
                    rbuf[rpos + q] = L.rot[2 * (m - 1 - (q >> 1)) + (q & 1)];
                if (lane == 0) {
                    rbuf[rpos + 2L * cnt] = (double)ilo;
                    rbuf[rpos + 2L * cnt + 1] = (double)m;
                }
                rpos += 2L * cnt + 2;
            } else {
                ovf = true;
            }
            if (lane < nvec) {
                double* z = V.at(lane);
                double z1 = z[m];
                for (int ii = m - 1; ii >= ilo; --ii) {
                    const double cc = L.rot[2 * ii], ss = L.rot[2 * ii + 1];
                    const double z0 = z[ii];
                    z[ii + 1] = fma(ss, z0, cc * z1);
                    z1 = fma(cc, z0, -ss * z1);
                }
                z[ilo] = z1;
            }
            if (!split) {
                d[l] = d[l] - p;
                e[l] = g;
                e[m] = 0.0;
            }
            __syncthreads();
        }
    }
    return 0;
}

// y <- S y: the recorded sweeps in reverse order, each sweep's rotations in
// increasing index (the inverse of the forward update)
__device__ void ql_back_vec(EigLds& L, const VecSet& V, int nvec, const double* rbuf, long rpos,
                            int lane)
{
    long pos = rpos;
    while (pos > 0) {
        const int ilo = (int)rbuf[pos - 2], m = (int)rbuf[pos - 1];
        const int cnt = m - ilo;
        const long base = pos - 2 - 2L * cnt;
        for (int q = lane; q < 2 * cnt; q += 64)
            L.rot[2 * (m - 1 - (q >> 1)) + (q & 1)] = rbuf[base + q];
        __syncthreads();
        if (lane < nvec) {
            double* z = V.at(lane);
            double zl = z[ilo];
            for (int ii = ilo; ii < m; ++ii) {
                const double cc = L.rot[2 * ii], ss = L.rot[2 * ii + 1];
                const double zi1 = z[ii + 1];
                z[ii] = fma(cc, zl, ss * zi1);
                zl = fma(-ss, zl, cc * zi1);
            }
            z[m] = zl;
        }
        __syncthreads();
        pos = base;
    }
}

// one point of the large-k vector path (no trans/pao outputs)
__device__ void big_vec_point(const LekfJob& J, int b, double* rec, double* rbuf, long rcap,
                              EigLds& L, int lane)
{
    const int k = J.k, lda = k;
    const long off = J.loc_off[b];
    const int p = J.loc_cnt ? J.loc_cnt[b] : (int)(J.loc_off[b + 1] - off);
    double alpha = J.infl_in ? J.infl_in[b] : J.infl_fixed;
    int status = 0;
    const double sa = sqrt(alpha);
    const bool apply = (J.flags & LEKF_F_APPLY) != 0;
    const int lev = apply ? b / J.nij : 0, ij = apply ? b - lev * J.nij : 0;
    const int n3 = apply ? J.nv3d : 0, n2 = (apply && lev == 0) ? J.nv2d : 0;
    const int nrow = n3 + n2;
    const int ks = k + 1;  // row stride of R: odd, so lane-per-row LDS accesses spread over banks
    if (nrow > 0) {
        for (int w = lane; w < nrow * k; w += 64) {
            const int r = w / k, m = w - r * k;
            const double* src = (r < n3) ? J.x3b + ((size_t)b * n3 + r) * k
                                         : J.x2b + ((size_t)ij * J.nv2d + (r - n3)) * k;
            L.X[r * k + m] = src[m];
        }
        __syncthreads();
        for (int r = lane; r < nrow; r += 64) {
            double s = L.X[r * k];
            for (int m = 1; m < k; ++m) s += L.X[r * k + m];
            L.M[r] = s / (double)k;
        }
        __syncthreads();
        for (int w = lane; w < nrow * k; w += 64) L.X[w] -= L.M[w / k];
        __syncthreads();
    }
    if (p > 0) {
        for (int i = lane; i < k; i += 64) L.B[i] = rec[(size_t)k * k + i];
        const double p1 = rec[(size_t)k * k + k], p3 = rec[(size_t)k * k + k + 1];
        __syncthreads();
        double tr = 0.0;
        for (int i = 0; i < k; ++i) tr += rec[(size_t)i * lda + i];
        const double p2 = tr / (double)(k - 1);
        if ((J.flags & LEKF_F_MINFL) && J.minfl > 0.0 && alpha < J.minfl) alpha = J.minfl;
        const double rho = 1.0 / alpha;
        __syncthreads();
        for (int i = lane; i < k; i += 64) rec[(size_t)i * lda + i] += (double)(k - 1) * rho;
        for (int w = lane; w < nrow * k; w += 64) L.R[(w / k) * ks + w % k] = L.X[w];
        __syncthreads();
        big_tridiag_sym<4>(rec, k, lda, L.d, L.e, L.tau, L.rot, L.rot + k, L.U, lane);
        const VecSet Vf{L.R, nrow, ks, L.B, nullptr};
        big_refl_vecs<4>(rec, k, lda, L.tau, Vf, nrow + 1, true, lane);
        __syncthreads();
        long rpos = 0;
        bool ovf = false;
        if (eig_ql_vec(L, k, lane, Vf, nrow + 1, rbuf, rcap, rpos, ovf)) status |= LEKF_PT_NOCONV;
        if (ovf) status |= LEKF_PT_NOCONV;
        __syncthreads();
        double lmax = 0.0, lmin = 1.0e308;
        for (int j = 0; j < k; ++j) {
            const double lj = L.d[j];
            lmax = lj > lmax ? lj : lmax;
            lmin = lj < lmin ? lj : lmin;
        }
        if (!(lmax > 0.0) || lmin < fabs(lmax) * 1.4901161193847656e-08) status |= LEKF_PT_RANKDEF;
        // u = L^-1 V^T b;  X' w-bar = (V^T x')^T u;  rows <- sqrt((k-1)/L) V^T x'
        for (int j = lane; j < k; j += 64) L.U[j] = L.B[j] / L.d[j];
        __syncthreads();
        for (int r = lane; r < nrow; r += 64) {
            double s = 0.0;
            for (int j = 0; j < k; ++j) s = fma(L.R[r * ks + j], L.U[j], s);
            L.Cw[r] = s;
        }
        const bool want_wb = J.transm != nullptr;
        if (want_wb)
            for (int j = lane; j < k; j += 64) L.Wb[j] = L.U[j];
        __syncthreads();
        for (int w = lane; w < nrow * k; w += 64) {
            const int j = w % k;
            L.R[(w / k) * ks + j] *= sqrt((double)(k - 1) / L.d[j]);
        }
        __syncthreads();
        const VecSet Vb{L.R, nrow, ks, L.Wb, nullptr};
        const int nvb = nrow + (want_wb ? 1 : 0);
        if (!ovf) ql_back_vec(L, Vb, nvb, rbuf, rpos, lane);
        big_refl_vecs<4>(rec, k, lda, L.tau, Vb, nvb, false, lane);
        __syncthreads();
        if (want_wb)
            for (int i = lane; i < k; i += 64) J.transm[(size_t)b * k + i] = L.Wb[i];
        if (lane == 0 && J.infl_out)
            J.infl_out[b] = (J.flags & LEKF_F_FIXED_INFL) ? J.infl_fixed
                                                          : adaptive_alpha(alpha, p1, p2, p3);
    } else {
        if (J.transm)
            for (int i = lane; i < k; i += 64) J.transm[(size_t)b * k + i] = 0.0;
        if (lane == 0 && J.infl_out)
            J.infl_out[b] = (J.flags & LEKF_F_FIXED_INFL) ? J.infl_fixed : alpha;
    }
    if (lane == 0 && J.status) J.status[b] = status;
    if (nrow == 0) return;
    const uint32_t mask = J.upd_mask ? J.upd_mask[lev] : 0xffffffffu;
    for (int w = lane; w < nrow * k; w += 64) {
        const int r = w / k, m = w - r * k;
        const bool upd = (r >= n3) || ((mask >> r) & 1u);
        double a;
        if (!upd)
            a = L.M[r] + L.X[r * k + m];
        else if (p > 0)
            a = L.M[r] + (L.Cw[r] + L.R[r * ks + m]);
        else
            a = L.M[r] + sa * L.X[r * k + m];
        double* dst = (r < n3) ? J.x3a + ((size_t)b * n3 + r) * k
                               : J.x2a + ((size_t)ij * J.nv2d + (r - n3)) * k;
        dst[m] = a;
    }
}

// One point of the large-k dual form (0 < p < k, no trans/pao).  With
// B = diag(sqrt w) Y (p x k), G = B B^T = U S U^T (in the workspace record,
// p x p) and gamma = (k-1)/alpha, C = gamma I + B^T B has the eigenvalues
// gamma + s_j (eigenvectors B^T u_j / sqrt(s_j)) and gamma (k - p times), so
//   f(C) x = f(gamma) x + B^T U diag((f(gamma + s) - f(gamma)) / s) U^T B x
// for f(l) = sqrt((k-1)/l), and C^-1 B^T = B^T (gamma I + G)^-1 gives
//   w-bar = B^T U diag(1/(gamma + s)) U^T (sqrt(w) d).
// The p x p eigenproblem runs exactly as the k x k one of big_vec_point,
// carrying sqrt(w) d and B x'_r through Q^T and the QL rotations.
__device__ void big_dual_point(const LekfJob& J, int b, double* rec, double* rbuf, long rcap,
                               EigLds& L, int lane)
{
    const int k = J.k;
    const long off = J.loc_off[b];
    const int p = J.loc_cnt ? J.loc_cnt[b] : (int)(J.loc_off[b + 1] - off);
    double alpha = J.infl_in ? J.infl_in[b] : J.infl_fixed;
    int status = 0;
    const bool wloc = (J.flags & LEKF_F_WLOC) != 0;
    const bool apply = (J.flags & LEKF_F_APPLY) != 0;
    const int lev = apply ? b / J.nij : 0, ij = apply ? b - lev * J.nij : 0;
    const int n3 = apply ? J.nv3d : 0, n2 = (apply && lev == 0) ? J.nv2d : 0;
    const int nrow = n3 + n2;
    const int ks = k + 1;
    if (nrow > 0) {
        for (int w = lane; w < nrow * k; w += 64) {
            const int r = w / k, m = w - r * k;
            const double* src = (r < n3) ? J.x3b + ((size_t)b * n3 + r) * k
                                         : J.x2b + ((size_t)ij * J.nv2d + (r - n3)) * k;
            L.X[r * k + m] = src[m];
        }
        __syncthreads();
        for (int r = lane; r < nrow; r += 64) {
            double s = L.X[r * k];
            for (int m = 1; m < k; ++m) s += L.X[r * k + m];
            L.M[r] = s / (double)k;
        }
        __syncthreads();
        for (int w = lane; w < nrow * k; w += 64) L.X[w] -= L.M[w / k];
        __syncthreads();
    }
    // weights, sqrt(w) d, P1 = sum w d^2, P3 = sum rho (common_letkf.f90:219-227)
    double p1 = 0.0, p3 = 0.0;
    for (int i = lane; i < p; i += 64) {
        const long o = J.loc_idx[off + i];
        const double rho = J.loc_rho[off + i];
        const double rd = J.loc_rdiag ? J.loc_rdiag[off + i] : J.rdiag[o];
        const double w = (wloc ? 1.0 : rho) / rd;
        const double dd = J.dep[o];
        L.sw[i] = sqrt(w);
        L.B[i] = L.sw[i] * dd;
        p1 = fma(w * dd, dd, p1);
        p3 += rho;
    }
    p1 = wave_sum(p1);
    p3 = wave_sum(p3);
    double tr = 0.0;
    for (int i = 0; i < p; ++i) tr += rec[(size_t)i * p + i];
    const double p2 = tr / (double)(k - 1);
    if ((J.flags & LEKF_F_MINFL) && J.minfl > 0.0 && alpha < J.minfl) alpha = J.minfl;
    const double gam = (double)(k - 1) / alpha;
    __syncthreads();
    // y_r = B x'_r: one local obs row at a time, lanes over members
    for (int i = 0; i < p; ++i) {
        const long o = J.loc_idx[off + i];
        double y[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int m = lane + 64 * t;
            y[t] = m < k ? J.hdxf[o * J.y_so + (long)m * J.y_sm] : 0.0;
        }
        for (int r = 0; r < nrow; ++r) {
            double s = 0.0;
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (lane + 64 * t < k) s = fma(y[t], L.X[r * k + lane + 64 * t], s);
            s = wave_sum(s);
            if (lane == 0) L.R[r * ks + i] = L.sw[i] * s;
        }
    }
    __syncthreads();
    big_tridiag_sym<4>(rec, p, p, L.d, L.e, L.tau, L.rot, L.rot + k, L.U, lane);
    const VecSet Vf{L.R, nrow, ks, L.B, nullptr};
    big_refl_vecs<4>(rec, p, p, L.tau, Vf, nrow + 1, true, lane);
    __syncthreads();
    long rpos = 0;
    bool ovf = false;
    if (eig_ql_vec(L, p, lane, Vf, nrow + 1, rbuf, rcap, rpos, ovf)) status |= LEKF_PT_NOCONV;
    if (ovf) status |= LEKF_PT_NOCONV;
    __syncthreads();
    // eigenvalues of C: gamma + s_j and gamma; truncation test (common_mtx.f90:66-74)
    double smax = L.d[0], smin = L.d[0];
    for (int j = 1; j < p; ++j) {
        smax = L.d[j] > smax ? L.d[j] : smax;
        smin = L.d[j] < smin ? L.d[j] : smin;
    }
    const double lmax = gam + (smax > 0.0 ? smax : 0.0);
    const double lmin = smin < 0.0 ? gam + smin : gam;
    if (!(lmax > 0.0) || lmin < fabs(lmax) * 1.4901161193847656e-08) status |= LEKF_PT_RANKDEF;
    // c_b = (gamma + s)^-1 U^T sqrt(w) d;  X' w-bar = (U^T B x')^T c_b
    for (int j = lane; j < p; j += 64) L.U[j] = L.B[j] / (gam + L.d[j]);
    __syncthreads();
    for (int r = lane; r < nrow; r += 64) {
        double s = 0.0;
        for (int j = 0; j < p; ++j) s = fma(L.R[r * ks + j], L.U[j], s);
        L.Cw[r] = s;
    }
    const bool want_wb = J.transm != nullptr;
    if (want_wb)
        for (int j = lane; j < p; j += 64) L.Wb[j] = L.U[j];
    __syncthreads();
    // (f(gamma + s) - f(gamma)) / s without cancellation
    const double sg = sqrt(gam), skm1 = sqrt((double)(k - 1));
    for (int w = lane; w < nrow * p; w += 64) {
        const int r = w / p, j = w - r * p;
        const double sgs = sqrt(gam + L.d[j]);
        L.R[r * ks + j] *= -skm1 / (sg * sgs * (sg + sgs));
    }
    __syncthreads();
    const VecSet Vb{L.R, nrow, ks, L.Wb, nullptr};
    const int nvb = nrow + (want_wb ? 1 : 0);
    if (!ovf) ql_back_vec(L, Vb, nvb, rbuf, rpos, lane);
    big_refl_vecs<4>(rec, p, p, L.tau, Vb, nvb, false, lane);
    __syncthreads();
    // coefficients of the local obs rows: sqrt(w_i) times the p-vectors
    for (int w = lane; w < nrow * p; w += 64) {
        const int r = w / p, i = w - r * p;
        L.R[r * ks + i] *= L.sw[i];
    }
    if (want_wb)
        for (int i = lane; i < p; i += 64) L.Wb[i] *= L.sw[i];
    __syncthreads();
    if (want_wb) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < p; ++i) {
            const long o = J.loc_idx[off + i];
            const double c = L.Wb[i];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int m = lane + 64 * t;
                if (m < k) acc[t] = fma(c, J.hdxf[o * J.y_so + (long)m * J.y_sm], acc[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (lane + 64 * t < k) J.transm[(size_t)b * k + lane + 64 * t] = acc[t];
    }
    if (lane == 0 && J.infl_out)
        J.infl_out[b] = (J.flags & LEKF_F_FIXED_INFL) ? J.infl_fixed
                                                      : adaptive_alpha(alpha, p1, p2, p3);
    if (lane == 0 && J.status) J.status[b] = status;
    if (nrow == 0) return;
    const uint32_t mask = J.upd_mask ? J.upd_mask[lev] : 0xffffffffu;
    const double fg = sqrt((double)(k - 1) / gam);
    for (int r = 0; r < nrow; ++r) {
        const bool upd = (r >= n3) || ((mask >> r) & 1u);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        if (upd)
            for (int i = 0; i < p; ++i) {
                const long o = J.loc_idx[off + i];
                const double c = L.R[r * ks + i];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int m = lane + 64 * t;
                    if (m < k) acc[t] = fma(c, J.hdxf[o * J.y_so + (long)m * J.y_sm], acc[t]);
                }
            }
        double* dst = (r < n3) ? J.x3a + ((size_t)b * n3 + r) * k
                               : J.x2a + ((size_t)ij * J.nv2d + (r - n3)) * k;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int m = lane + 64 * t;
            if (m < k) {
                const double xp = L.X[r * k + m];
                dst[m] = upd ? L.M[r] + (L.Cw[r] + fma(fg, xp, acc[t])) : L.M[r] + xp;
            }
        }
    }
}

// One wave per point: C and sums from the Gram workspace, eigendecomposition
// (KR > 0: in registers for k <= KR; KR = 0: in LDS; BIG > 0: in the point's
// workspace record in global memory, BIG rows per lane), W assembly, inflation,
// the optional trans/transm/pao outputs and the fused apply
// Xa = x-mean + X' T evaluated as x-mean + (X' V) diag(sqrt((k-1)/L)) V^T
// + (X' w-bar) 1^T.
// The split eigensolver (32 < k <= 64, three kernels per sub-batch): the
// point's record carries, after the Gram, the Householder Q (transposed: Q_rt
// at qoff + t k + r), the tridiagonal d and e, then the eigenvalues, the
// rotation-stream length and the QL status written by the QL kernel.  The
// rotation stream of a point holds, per QL sweep, a header (m, count) and the
// sweep's (c, s) pairs for i = m-1 down to m-count.
struct SplitRec {
    long qoff, doff, eoff, lamoff, cntoff, flagoff, qroff;
    const double* rot;  // this point's rotation stream
    long rcap;          // its capacity in (c, s) entries
};
constexpr int kSplitOverflow = 8;  // flag: the stream overflowed (the point is solved again in full)

// Flat point of slot `slot` of a split kernel: pass 0 the sub-batch's own
// (gp0 + slot, nb slots), pass 1 the redo list's (rlist[0] entries from
// rlist[1], of which the first nb were written: the count goes on past the
// list's capacity); false past the end.
__device__ inline bool split_slot_point(long slot, long gp0, long nb, int pass,
                                        const int* rlist, long& gp)
{
    if (pass) {
        if (slot >= nb || slot >= rlist[0]) return false;
        gp = rlist[1 + slot];
        return true;
    }
    gp = gp0 + slot;
    return slot < nb;
}

// Whether a split kernel of pass `pass` (0: the launch's own precision, 1: the
// FP64 redo) leaves a point with local obs to the other pass.
__device__ inline bool split_other_pass(const double* rec, long flagoff, int pass)
{
    return ((((int)rec[flagoff]) & kSplitRedo) != 0) != (pass != 0);
}

__host__ __device__ inline long split_qoff(int kmax) { return gram_stride(kmax); }
__host__ __device__ inline long split_stride(int kmax)
{
    return (gram_stride(kmax) + 2L * kmax * kmax + 3L * kmax + 2 + 1) & ~1L;
}
__host__ __device__ inline long split_rcap(int kmax) { return 2L * kmax * kmax + 8L * kmax; }
// per point: the column of each rotation-stream entry (-1: a sweep header),
// after the streams, in doubles
// shorts per point of the column stream, a multiple of 4 (8-byte rows)
__host__ __device__ inline long split_ridx_stride(int kmax) { return (split_rcap(kmax) + 3) & ~3L; }
__host__ __device__ inline long split_ridx_doubles(int kmax) { return split_ridx_stride(kmax) / 4; }

__host__ __device__ inline SplitRec split_rec(int kmax, const double* rot)
{
    SplitRec S;
    S.qoff = split_qoff(kmax);
    S.doff = S.qoff + (long)kmax * kmax;
    S.eoff = S.doff + kmax;
    S.lamoff = S.eoff + kmax;
    S.cntoff = S.lamoff + kmax;
    S.flagoff = S.cntoff + 1;
    S.qroff = S.flagoff + 1;  // Q again, row-major (Q[i][j] at qroff + i * kmax + j)
    S.rot = rot;
    S.rcap = split_rcap(kmax);
    return S;
}

// Eigenvectors from the split solve: Q into Z in LDS (lane r: row r), the
// recorded QL rotations replayed in order on each lane's row (the same
// operations ql_reg performs as it goes, so the result is the same bit for
// bit; a sweep costs its own length), then the eigenvalues to LDS.
template <int KR>
__device__ void eig_replay(EigLds& L, int n, int lda, int lane, const double* rec,
                           const SplitRec& S)
{
    for (int t = 0; t < n; ++t)
        if (lane < n) L.Z[lane * lda + t] = rec[S.qoff + (long)t * n + lane];
    const long ntot = (long)rec[S.cntoff];
    const double* rs = S.rot;
    double* zr = L.Z + lane * lda;
    long pos = 0;
    while (pos < ntot) {
        const int m = (int)rs[2 * pos], cnt = (int)rs[2 * pos + 1];
        ++pos;
        for (int j = lane; j < cnt; j += 64) {
            L.rot[2 * j] = rs[2 * (pos + j)];
            L.rot[2 * j + 1] = rs[2 * (pos + j) + 1];
        }
        __syncthreads();
        if (lane < n && cnt > 0) {
            // column i+1 takes its final value at step i; column i carries on
            // (four steps at a time: their loads first, then the chain)
            double h = zr[m];
            int j = 0;
            for (; j + 4 <= cnt; j += 4) {
                double qi[4], c[4], sn[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    qi[u] = zr[m - 1 - (j + u)];
                    c[u] = L.rot[2 * (j + u)];
                    sn[u] = L.rot[2 * (j + u) + 1];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    zr[m - (j + u)] = fma(sn[u], qi[u], c[u] * h);
                    h = fma(c[u], qi[u], -sn[u] * h);
                }
            }
            for (; j < cnt; ++j) {
                const int i = m - 1 - j;
                const double c = L.rot[2 * j], sn = L.rot[2 * j + 1];
                const double qi = zr[i];
                zr[i + 1] = fma(sn, qi, c * h);
                h = fma(c, qi, -sn * h);
            }
            zr[m - cnt] = h;
        }
        __syncthreads();
        pos += cnt;
    }
    if (lane < n) L.d[lane] = rec[S.lamoff + lane];
    __syncthreads();
}

template <int KR, int BIG, bool REPLAY = false>
__device__ void eig_point(const LekfJob& J, int b, double* const rec, long gstride, int nvmax,
                          long rot_off, double* smem, int lane, const SplitRec* SR = nullptr)
{
    const int k = J.k, lda = BIG ? k : eig_lda(k);
    EigLds L = eig_carve(smem, k, nvmax, BIG ? rec : nullptr);
    if constexpr (BIG > 0) {
        if (!J.trans && !J.pao) {
            const long o0 = J.loc_off[b];
            const int pp = J.loc_cnt ? J.loc_cnt[b] : (int)(J.loc_off[b + 1] - o0);
            if (big_dual(J, pp, k))
                big_dual_point(J, b, rec, rec + rot_off, gstride - rot_off, L, lane);
            else
                big_vec_point(J, b, rec, rec + rot_off, gstride - rot_off, L, lane);
            return;
        }
    }
    // Z_ij (component i of eigenvector j once the eigendecomposition is done)
    auto Zv = [&](int i, int j) -> double {
        return BIG ? L.Z[i + (size_t)j * lda] : L.Z[i * lda + j];
    };
    const long off = J.loc_off[b];
    const int p = J.loc_cnt ? J.loc_cnt[b] : (int)(J.loc_off[b + 1] - off);
    double alpha = J.infl_in ? J.infl_in[b] : J.infl_fixed;
    const bool transm_mode = (J.flags & LEKF_F_TRANSM) != 0;
    int status = 0;
    const double sa = sqrt(alpha);  // p = 0 branch: T = sqrt(alpha) I

    if (p > 0) {
        const double* g = rec;
        // the split solve's result: rotations replayed; a point whose stream
        // overflowed is left to lekf_eig_fix_kernel
        const bool replay = REPLAY && !(((int)rec[SR->flagoff]) & kSplitOverflow);
        if (REPLAY && !replay) return;
        if (!BIG && !replay)
            for (int t = lane; t < k * k; t += 64) {
                const int i = t / k, j = t - i * k;
                L.Z[i * lda + j] = g[t];
            }
        for (int i = lane; i < k; i += 64) L.B[i] = g[(size_t)k * k + i];
        const double p1 = g[(size_t)k * k + k], p3 = g[(size_t)k * k + k + 1];
        __syncthreads();
        // P2 = trace(C) / (k-1), summed in index order as the reference does
        double tr = 0.0;
        if (replay)
            for (int i = 0; i < k; ++i) tr += g[(size_t)i * k + i];
        else
            for (int i = 0; i < k; ++i) tr += L.Z[i * lda + i];
        const double p2 = tr / (double)(k - 1);
        if ((J.flags & LEKF_F_MINFL) && J.minfl > 0.0 && alpha < J.minfl) alpha = J.minfl;
        const double rho = 1.0 / alpha;
        __syncthreads();
        if (!replay)
            for (int i = lane; i < k; i += 64) L.Z[i * lda + i] += (double)(k - 1) * rho;
        __syncthreads();
        if (replay) {
            if constexpr (KR > 0) eig_replay<KR>(L, k, lda, lane, rec, *SR);
            if (((int)rec[SR->flagoff]) & 1) status |= LEKF_PT_NOCONV;
        } else if constexpr (BIG > 0) {
            big_tridiag<(BIG > 0 ? BIG : 1)>(L.Z, k, lda, L.d, L.e, L.tau, lane);
            if (eig_ql<BIG>(L, k, lda, lane)) status |= LEKF_PT_NOCONV;
        } else if constexpr (KR > 0 && !REPLAY) {
            if (eig_reg<(KR > 0 ? KR : 4)>(L, k, lda, lane)) status |= LEKF_PT_NOCONV;
        } else {
            eig_tridiag(L, k, lda, lane);
            if (eig_ql(L, k, lda, lane)) status |= LEKF_PT_NOCONV;
        }
        __syncthreads();
        // reference truncation test (common_mtx.f90:66-74)
        double lmax = 0.0, lmin = 1.0e308;
        for (int j = 0; j < k; ++j) {
            const double lj = L.d[j];
            lmax = lj > lmax ? lj : lmax;
            lmin = lj < lmin ? lj : lmin;
        }
        if (!(lmax > 0.0) || lmin < fabs(lmax) * 1.4901161193847656e-08) status |= LEKF_PT_RANKDEF;
        // u = L^-1 V^T b ; w-bar = V u ; U <- sqrt((k-1)/lambda)
        for (int j = lane; j < k; j += 64) {
            double s = 0.0;
            for (int i = 0; i < k; ++i) s = fma(Zv(i, j), L.B[i], s);
            L.U[j] = s / L.d[j];
        }
        __syncthreads();
        for (int i = lane; i < k; i += 64) {
            double s = 0.0;
            for (int j = 0; j < k; ++j) s = fma(Zv(i, j), L.U[j], s);
            L.Wb[i] = s;
        }
        __syncthreads();
        for (int j = lane; j < k; j += 64) L.U[j] = sqrt((double)(k - 1) / L.d[j]);
        __syncthreads();
        if (J.trans) {
            double* tr = J.trans + (size_t)b * k * k;
            for (int w = lane; w < k * k; w += 64) {
                const int m = w / k, i = w - m * k;  // column-major, coalesced over i
                double s = 0.0;
                for (int l = 0; l < k; ++l) s = fma(Zv(i, l) * L.U[l], Zv(m, l), s);
                if (!transm_mode) s += L.Wb[i];
                tr[i + (size_t)m * k] = s;
            }
        }
        if (J.pao) {
            double* pa = J.pao + (size_t)b * k * k;
            for (int w = lane; w < k * k; w += 64) {
                const int m = w / k, i = w - m * k;
                double s = 0.0;
                for (int l = 0; l < k; ++l) s = fma(Zv(i, l) / L.d[l], Zv(m, l), s);
                pa[i + (size_t)m * k] = s;
            }
        }
        if (J.transm)
            for (int i = lane; i < k; i += 64) J.transm[(size_t)b * k + i] = L.Wb[i];
        if (lane == 0 && J.infl_out)
            J.infl_out[b] = (J.flags & LEKF_F_FIXED_INFL) ? J.infl_fixed
                                                          : adaptive_alpha(alpha, p1, p2, p3);
    } else {
        // no local obs (common_letkf.f90:82-96)
        if (J.trans) {
            double* tr = J.trans + (size_t)b * k * k;
            for (int w = lane; w < k * k; w += 64) {
                const int m = w / k, i = w - m * k;
                tr[i + (size_t)m * k] = (i == m) ? sa : 0.0;
            }
        }
        if (J.pao) {
            double* pa = J.pao + (size_t)b * k * k;
            for (int w = lane; w < k * k; w += 64) {
                const int m = w / k, i = w - m * k;
                pa[i + (size_t)m * k] = (i == m) ? alpha / (double)(k - 1) : 0.0;
            }
        }
        if (J.transm)
            for (int i = lane; i < k; i += 64) J.transm[(size_t)b * k + i] = 0.0;
        if (lane == 0 && J.infl_out)
            J.infl_out[b] = (J.flags & LEKF_F_FIXED_INFL) ? J.infl_fixed : alpha;
    }
    if (lane == 0 && J.status) J.status[b] = status;
    if (!(J.flags & LEKF_F_APPLY)) return;
    // fused apply: rows of the point (3-D at every level, 2-D at level 0)
    const int lev = b / J.nij, ij = b - lev * J.nij;
    const int n3 = J.nv3d, n2 = (lev == 0) ? J.nv2d : 0;
    const int nrow = n3 + n2;
    const uint32_t mask = J.upd_mask ? J.upd_mask[lev] : 0xffffffffu;
    for (int w = lane; w < nrow * k; w += 64) {
        const int r = w / k, m = w - r * k;
        const double* src = (r < n3) ? J.x3b + ((size_t)b * n3 + r) * k
                                     : J.x2b + ((size_t)ij * J.nv2d + (r - n3)) * k;
        L.X[r * k + m] = src[m];
    }
    __syncthreads();
    // ensemble mean in member order (speedy/common/common_speedy.f90:335-370)
    for (int r = lane; r < nrow; r += 64) {
        double s = L.X[r * k];
        for (int m = 1; m < k; ++m) s += L.X[r * k + m];
        L.M[r] = s / (double)k;
    }
    __syncthreads();
    for (int w = lane; w < nrow * k; w += 64) L.X[w] -= L.M[w / k];
    __syncthreads();
    if (p > 0) {
        // R = (X' V) diag(U), Cw = X' w-bar
        for (int w = lane; w < nrow * k; w += 64) {
            const int r = w / k, l = w - r * k;
            double s = 0.0;
            for (int j = 0; j < k; ++j) s = fma(L.X[r * k + j], Zv(j, l), s);
            L.R[r * k + l] = s * L.U[l];
        }
        for (int r = lane; r < nrow; r += 64) {
            double s = 0.0;
            for (int j = 0; j < k; ++j) s = fma(L.X[r * k + j], L.Wb[j], s);
            L.Cw[r] = s;
        }
        __syncthreads();
    }
    for (int w = lane; w < nrow * k; w += 64) {
        const int r = w / k, m = w - r * k;
        const bool upd = (r >= n3) || ((mask >> r) & 1u);
        double a;
        if (!upd) {
            a = L.M[r] + L.X[r * k + m];
        } else if (p > 0) {
            double s = 0.0;
            for (int l = 0; l < k; ++l) s = fma(L.R[r * k + l], Zv(m, l), s);
            a = L.M[r] + (L.Cw[r] + s);
        } else {
            a = L.M[r] + sa * L.X[r * k + m];
        }
        double* dst = (r < n3) ? J.x3a + ((size_t)b * n3 + r) * k
                               : J.x2a + ((size_t)ij * J.nv2d + (r - n3)) * k;
        dst[m] = a;
    }
}

// Waves per SIMD asked of the register eigen kernel: at KR = 40..48 the
// register QL fits two waves per SIMD with a few spills, which hides more of
// its latency than the spills cost (measured on the GFS-sized problem: the
// solve 28% faster at k = 40, 9.5% at 44, 2% at 48; at KR = 36 it is two
// waves without the hint and the hint's spills made it 2% slower; at KR = 52
// it spills 110 registers and the radar-sized cycle is 15% slower).
constexpr int eig_waves_per_simd(int kr) { return (kr >= 40 && kr <= 48) ? 2 : 1; }

template <int KR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(eig_waves_per_simd(KR))))
void lekf_eig_kernel(const LekfJob* __restrict__ jobs, int njobs, long gp0,
                     double* __restrict__ ws, long gstride, int nvmax)
{
    extern __shared__ double smem[];
    const long gp = gp0 + blockIdx.x;
    const LekfJob& J = jobs[lekf_job_of(jobs, njobs, gp)];
    eig_point<KR, 0>(J, (int)(gp - J.pt_begin + J.pt0), ws + (gp - gp0) * gstride, gstride, nvmax,
                     0L, smem, threadIdx.x);
}

// --- split eigensolver, 32 < k <= 64 --------------------------------------
// Kernel 1 (one wave per point): C + (k-1)/alpha I into LDS and the
// Householder tridiagonalisation (eig_tridiag64, as the fused kernel does);
// writes Q (transposed), d and e to the point's record.
__global__ __launch_bounds__(64) void lekf_tri_kernel(const LekfJob* __restrict__ jobs, int njobs,
                                                       long gp0, double* __restrict__ ws,
                                                       long gstride, int kmax, int pass,
                                                       int* __restrict__ rlist, int zero)
{
    extern __shared__ double smem[];
    // the launch's first sub-batch empties the redo list (before its QL kernel)
    if (zero && blockIdx.x == 0 && threadIdx.x == 0) rlist[0] = 0;
    long gp;
    if (!split_slot_point(blockIdx.x, gp0, (long)gridDim.x, pass, rlist, gp)) return;
    const LekfJob& J = jobs[lekf_job_of(jobs, njobs, gp)];
    const int b = (int)(gp - J.pt_begin + J.pt0);
    const int k = J.k, lane = threadIdx.x, lda = eig_lda(k);
    const int p = J.loc_cnt ? J.loc_cnt[b] : (int)(J.loc_off[b + 1] - J.loc_off[b]);
    if (p <= 0) return;
    double* rec = ws + (long)blockIdx.x * gstride;
    const SplitRec S = split_rec(kmax, nullptr);
    EigLds L = eig_carve_tri(smem, k);
    double alpha = J.infl_in ? J.infl_in[b] : J.infl_fixed;
    if ((J.flags & LEKF_F_MINFL) && J.minfl > 0.0 && alpha < J.minfl) alpha = J.minfl;
    const double shift = (double)(k - 1) * (1.0 / alpha);
    for (int t = lane; t < k * k; t += 64) {
        const int i = t / k, j = t - i * k;
        L.Z[i * lda + j] = rec[t];
    }
    __syncthreads();
    for (int i = lane; i < k; i += 64) L.Z[i * lda + i] += shift;
    __syncthreads();
    eig_tridiag64(L, k, lda, lane);
    if (lane < k) {
        // Z[r * lda + c] = Q[r][c]; both copies are written a row at a time
        for (int t = 0; t < k; ++t) {
            rec[S.qoff + (long)t * k + lane] = L.Z[lane * lda + t];   // Q[lane][t], column-major
            rec[S.qroff + (long)t * k + lane] = L.Z[t * lda + lane];  // Q[t][lane], row-major
        }
        rec[S.doff + lane] = L.d[lane];
        rec[S.eoff + lane] = L.e[lane];
    }
}

// Kernel 2 (one lane per point, 64 points per wave): implicit QL on the
// point's tridiagonal held in LDS (d_i at i*64 + lane, e_i after the KR rows
// of d), the same recurrence as ql_reg; each rotation (c, s) goes to the
// point's stream instead of rotating Q.  Writes the eigenvalues, the stream
// length and the flags (1: an eigenvalue did not converge, kSplitOverflow:
// the stream was too short) to the record.
template <int KR>
__global__ __launch_bounds__(64) void lekf_ql_kernel(const LekfJob* __restrict__ jobs, int njobs,
                                                      long gp0, long nb, double* __restrict__ ws,
                                                      long gstride, double* __restrict__ rot,
                                                      short* __restrict__ ridx, int kmax,
                                                      long rlim, int pass, double kthr,
                                                      int* __restrict__ rlist, long lcap)
{
    extern __shared__ double sm[];
    const int lane = threadIdx.x;
    const long pt = (long)blockIdx.x * 64 + lane;  // the slot
    long gp;
    if (!split_slot_point(pt, gp0, nb, pass, rlist, gp)) return;  // no barrier below
    const LekfJob& J = jobs[lekf_job_of(jobs, njobs, gp)];
    const int b = (int)(gp - J.pt_begin + J.pt0);
    const int n = J.k;
    const int p = J.loc_cnt ? J.loc_cnt[b] : (int)(J.loc_off[b + 1] - J.loc_off[b]);
    if (p <= 0) return;
    double* rec = ws + pt * gstride;
    // a stream holds up to rlim rotations (at most its stride split_rcap)
    const long stride = split_rcap(kmax);
    const long rcap = rlim > 0 && rlim < stride ? rlim : stride;
    double* rs = rot + pt * 2 * stride;
    short* rx = ridx + pt * split_ridx_stride(kmax);
    const SplitRec S = split_rec(kmax, nullptr);
    double* d = sm + lane;
    double* e = sm + (long)KR * 64 + lane;
    for (int i = 0; i < n; ++i) {
        d[i * 64] = rec[S.doff + i];
        e[i * 64] = rec[S.eoff + i];
    }
    long nr = 0;
    int bad = 0;
    double tst = 0.0;
    for (int l = 0; l < n && !bad; ++l) {
        tst = fmax(tst, fabs(d[l * 64]) + fabs(e[l * 64]));
        for (int iter = 0;; ++iter) {
            int m = l;
            while (m < n - 1 && !(tst + fabs(e[m * 64]) == tst)) ++m;
            if (m == l) break;
            if (iter == kQlMaxIter) {
                bad = 1;
                break;
            }
            const double dl = d[l * 64], el = e[l * 64];
            double g = (d[(l + 1) * 64] - dl) / (2.0 * el);
            double r = fabs(g) > 1.0e100 ? fabs(g) : sqrt(fma(g, g, 1.0));
            g = d[m * 64] - dl + el / (g + (g >= 0.0 ? r : -r));
            double sn = 1.0, c = 1.0, pp = 0.0, di1 = d[m * 64];
            const long hpos = nr++;
            int cnt = 0;
            bool split = false;
            for (int i = m - 1; i >= l; --i) {
                const double ei = e[i * 64], di = d[i * 64];
                const double f = sn * ei, bb = c * ei;
                const double x = fma(f, f, g * g);
                if (x == 0.0) {
                    d[(i + 1) * 64] = di1 - pp;
                    e[m * 64] = 0.0;
                    split = true;
                    break;
                }
                const double y = rsq_nr(x);
                r = x * y;
                sn = f * y;
                c = g * y;
                e[(i + 1) * 64] = r;
                g = di1 - pp;
                r = fma(di - g, sn, 2.0 * c * bb);
                pp = sn * r;
                d[(i + 1) * 64] = g + pp;
                g = fma(c, r, -bb);
                di1 = di;
                if (nr < rcap) {
                    *reinterpret_cast<double2*>(rs + 2 * nr) = make_double2(c, sn);
                    rx[nr] = (short)i;
                }
                ++nr;
                ++cnt;
            }
            if (!split) {
                d[l * 64] -= pp;
                e[l * 64] = g;
                e[m * 64] = 0.0;
            }
            if (hpos < rcap) {
                *reinterpret_cast<double2*>(rs + 2 * hpos) = make_double2((double)m, (double)cnt);
                rx[hpos] = -1;
            }
        }
    }
    double lmin = d[0], lmax = d[0];
    for (int i = 0; i < n; ++i) {
        rec[S.lamoff + i] = d[i * 64];
        lmin = fmin(lmin, d[i * 64]);
        lmax = fmax(lmax, d[i * 64]);
    }
    // pass 0 of a mixed launch (kthr > 0): a point whose C + (k-1)/alpha I has
    // condition number above kthr goes on the redo list (while it has room: a
    // point past lcap keeps this pass's solution and no FP64 status), and its
    // flag tells this pass's later kernels to leave it; the FP64 pass's flag
    // keeps the bit so that its own later kernels find the point
    int redo = pass ? kSplitRedo : 0;
    if (!pass && kthr > 0.0 && !bad && J.hdxf && lmin > 0.0 && lmax > kthr * lmin) {
        const int pos = atomicAdd(rlist, 1);
        if (pos < lcap) {
            rlist[1 + pos] = (int)gp;
            redo = kSplitRedo;
        }
    }
    rec[S.cntoff] = (double)(nr < rcap ? nr : rcap);
    rec[S.flagoff] = (double)(bad | (nr > rcap ? kSplitOverflow : 0) | redo);
}

// Kernel 3 for jobs that output trans or pao (one wave per point): the
// eigenvectors by replaying the rotations on Q, then the W assembly,
// inflation, outputs and the fused apply of eig_point.  A point whose stream overflowed is left to lekf_eig_fix_kernel.
template <int KR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(eig_waves_per_simd(KR))))
void lekf_eig3_kernel(const LekfJob* __restrict__ jobs, int njobs, long gp0,
                      double* __restrict__ ws, long gstride, int nvmax, int kmax,
                      const double* __restrict__ rot, int pass, const int* __restrict__ rlist)
{
    extern __shared__ double smem[];
    const long slot = blockIdx.x;
    long gp;
    if (!split_slot_point(slot, gp0, (long)gridDim.x, pass, rlist, gp)) return;
    const LekfJob& J = jobs[lekf_job_of(jobs, njobs, gp)];
    if (!J.trans && !J.pao) return;  // lekf_evf_kernel's
    const int b = (int)(gp - J.pt_begin + J.pt0);
    const int p = J.loc_cnt ? J.loc_cnt[b] : (int)(J.loc_off[b + 1] - J.loc_off[b]);
    double* rec = ws + slot * gstride;
    const SplitRec S = split_rec(kmax, rot + slot * 2 * split_rcap(kmax));
    // no local obs: pass 0 writes the outputs; else the pass the flag names
    if (p <= 0 ? pass != 0 : split_other_pass(rec, S.flagoff, pass)) return;
    eig_point<KR, 0, true>(J, b, rec, gstride, nvmax, 0L, smem, threadIdx.x, &S);
    if (pass && threadIdx.x == 0 && J.status) J.status[b] |= LEKF_PT_FP64;
}

// Kernel 3 for the other jobs (no trans or pao output): the analysis without
// forming the eigenvectors.  With A = Q T Q^T (kernel 1) and T = R L R^T (the
// recorded rotations, R = G_1 G_2 ...), V = Q R, and every quantity the apply
// needs is a row vector pushed through Q, R, a diagonal, R^T and Q^T:
//   u = L^-1 V^T b             : row b^T -> b^T Q -> (.) R -> / lambda
//   X'V diag(U) V^T            : row x   -> x Q   -> (.) R -> * U -> (.) R^T -> (.) Q^T
//   X' w-bar = (X'V) . u       : the row after R, dotted with u
//   w-bar = V u                : u -> (.) R^T -> (.) Q^T
// (U = sqrt((k-1)/lambda), common/common_letkf.f90:119-147).  A rotation costs
// 6 flops per row instead of 6k for the whole of V, so G = nvmax + 1 lanes
// serve one point (its state rows, then b) and a wave takes 64 / G points;
// each lane keeps its row in LDS and follows its own point's rotation
// stream (column of each entry from the QL kernel, -1 for a sweep header),
// forward for R and backward with the transposed rotations for R^T.
// Results agree with the explicit form to rounding.
constexpr int kEvfChunk = 24;  // rotation-stream entries loaded at once per lane

// Entries e0 .. e0 + kEvfChunk - 1 of a lane's rotation stream (e0 a multiple
// of 4): (c, s) pairs as 16-byte loads, columns four at a time (the column
// rows are 8-byte aligned and padded to a multiple of 4); entries from ntot on
// have column -1.
__device__ __forceinline__ void evf_chunk(double* cc, double* ss, int* ii, const double* rs,
                                          const short* rx, long e0, long ntot)
{
#pragma unroll
    for (int u = 0; u < kEvfChunk; u += 4) {
        const long e = e0 + u;
        short4 q = make_short4(-1, -1, -1, -1);
        if (e < ntot) q = *reinterpret_cast<const short4*>(rx + e);
        ii[u] = e < ntot ? q.x : -1;
        ii[u + 1] = e + 1 < ntot ? q.y : -1;
        ii[u + 2] = e + 2 < ntot ? q.z : -1;
        ii[u + 3] = e + 3 < ntot ? q.w : -1;
    }
#pragma unroll
    for (int u = 0; u < kEvfChunk; ++u) {
        const long e = e0 + u;
        double2 cs = make_double2(0.0, 0.0);
        if (e < ntot) cs = *reinterpret_cast<const double2*>(rs + 2 * e);
        cc[u] = cs.x;
        ss[u] = cs.y;
    }
}

// The rows of a point group times a k x k matrix M (row-major, global
// memory), in place in the rows: row_r <- row_r M for the group's G <= RM rows
// (LDS, stride ld).  Lane v of the group forms columns 8v .. 8v + 7 of every
// row (G * 8 >= k), so each element of M is loaded by one lane of the group
// and the loads are contiguous; each sum runs over l in order, as the
// one-row-per-lane form does.  Uniform call (barriers inside).
template <int RM>
__device__ __forceinline__ void evf_coop_mv(double* rows, int ld, int G, int k, const double* M,
                                            int v, bool solve)
{
    double acc[RM][8];
#pragma unroll
    for (int r = 0; r < RM; ++r)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[r][j] = 0.0;
    const int c0 = v * 8;
    const bool mine = solve && c0 < k;
    if (mine) {
        for (int l = 0; l < k; ++l) {
            const double* ml = M + (size_t)l * k + c0;
            double q[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) q[j] = (c0 + j < k) ? ml[j] : 0.0;
#pragma unroll
            for (int r = 0; r < RM; ++r)
                if (r < G) {
                    const double xr = rows[(size_t)r * ld + l];
#pragma unroll
                    for (int j = 0; j < 8; ++j) acc[r][j] = fma(xr, q[j], acc[r][j]);
                }
        }
    }
    __syncthreads();
    if (mine)
#pragma unroll
        for (int r = 0; r < RM; ++r)
            if (r < G)
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (c0 + j < k) rows[(size_t)r * ld + c0 + j] = acc[r][j];
    __syncthreads();
}

template <int KR>
__global__ __launch_bounds__(64) void lekf_evf_kernel(const LekfJob* __restrict__ jobs, int njobs,
                                                       long gp0, long nb,
                                                       const double* __restrict__ ws, long gstride,
                                                       const double* __restrict__ rot,
                                                       const short* __restrict__ ridx, int nvmax,
                                                       int kmax, int pass,
                                                       const int* __restrict__ rlist)
{
    extern __shared__ double sm[];
    const int lane = threadIdx.x;
    const int G = nvmax + 1, P = 64 / G;
    const int grp = lane / G, v = lane - grp * G;
    const long pt = (long)blockIdx.x * P + grp;  // the slot
    const int ld = kmax | 1;
    double* row = sm + (size_t)lane * ld;
    const long rstride = split_rcap(kmax);
    const SplitRec S = split_rec(kmax, nullptr);

    // this lane's point and row
    const LekfJob* Jp = nullptr;
    int b = 0, k = 0, p = 0, nrow = 0, n3 = 0, lev = 0, ij = 0;
    bool work = false;  // a point this kernel solves
    long gp = 0;
    if (grp < P && split_slot_point(pt, gp0, nb, pass, rlist, gp)) {
        Jp = &jobs[lekf_job_of(jobs, njobs, gp)];
        b = (int)(gp - Jp->pt_begin + Jp->pt0);
        k = Jp->k;
        p = Jp->loc_cnt ? Jp->loc_cnt[b] : (int)(Jp->loc_off[b + 1] - Jp->loc_off[b]);
        work = !Jp->trans && !Jp->pao;
        // no local obs: pass 0 writes the outputs; else the pass the flag names
        if (work && (p <= 0 ? pass != 0 : split_other_pass(ws + pt * gstride, S.flagoff, pass)))
            work = false;
        if (work && p > 0 && (((int)ws[pt * gstride + S.flagoff]) & kSplitOverflow)) work = false;
        if ((Jp->flags & LEKF_F_APPLY)) {
            lev = b / Jp->nij;
            ij = b - lev * Jp->nij;
            n3 = Jp->nv3d;
            nrow = n3 + (lev == 0 ? Jp->nv2d : 0);
        }
    }
    const bool isb = work && v == nvmax;           // the lane of b
    const bool isx = work && v < nrow;             // a state row
    const double* rec = ws + pt * gstride;
    const double* xsrc = nullptr;
    double* xdst = nullptr;
    bool upd = true;
    if (isx) {
        const LekfJob& J = *Jp;
        xsrc = (v < n3) ? J.x3b + ((size_t)b * n3 + v) * k : J.x2b + ((size_t)ij * J.nv2d + (v - n3)) * k;
        xdst = (v < n3) ? J.x3a + ((size_t)b * n3 + v) * k : J.x2a + ((size_t)ij * J.nv2d + (v - n3)) * k;
        const uint32_t mask = J.upd_mask ? J.upd_mask[lev] : 0xffffffffu;
        upd = (v >= n3) || ((mask >> v) & 1u);
    }
    double alpha = 1.0;
    if (work) alpha = Jp->infl_in ? Jp->infl_in[b] : Jp->infl_fixed;
    // the row's mean in member order (speedy/common/common_speedy.f90:335-370)
    double mean = 0.0;
    if (isx) {
        double sx = xsrc[0];
        for (int m = 1; m < k; ++m) sx += xsrc[m];
        mean = sx / (double)k;
    }
    const bool solve = work && p > 0;
    if (work && p <= 0) {
        // no local obs (common_letkf.f90:82-96): T = sqrt(alpha) I
        const double sa = sqrt(alpha);
        if (isx)
            for (int m = 0; m < k; ++m) {
                const double xm = xsrc[m] - mean;
                xdst[m] = upd ? mean + sa * xm : mean + xm;
            }
        if (isb) {
            const LekfJob& J = *Jp;
            if (J.transm)
                for (int i = 0; i < k; ++i) J.transm[(size_t)b * k + i] = 0.0;
            if (J.infl_out) J.infl_out[b] = (J.flags & LEKF_F_FIXED_INFL) ? J.infl_fixed : alpha;
            if (J.status) J.status[b] = 0;
        }
    }
    const bool act = solve && (isx || isb);
    // the group's rows go through the matrix products together when they fit
    // (G <= 8 rows, 8 columns per lane); else each lane its own row
    const bool coop = G <= 8 && G * 8 >= kmax;
    // row <- x Q (Q row-major at qroff)
    if (act) {
        if (isx)
            for (int m = 0; m < k; ++m) row[m] = xsrc[m] - mean;
        else
            for (int i = 0; i < k; ++i) row[i] = rec[(size_t)k * k + i];
    }
    if (coop) {
        __syncthreads();
        evf_coop_mv<8>(sm + (size_t)grp * G * ld, ld, G, k, rec + S.qroff, v, solve);
    } else if (act) {
        double y[KR];
#pragma unroll
        for (int t = 0; t < KR; ++t) y[t] = 0.0;
        const double* qr = rec + S.qroff;
        for (int l = 0; l < k; ++l) {
            const double xl = row[l];
            const double* ql = qr + (size_t)l * k;
#pragma unroll
            for (int t = 0; t < KR; ++t)
                if (t < k) y[t] = fma(xl, ql[t], y[t]);
        }
#pragma unroll
        for (int t = 0; t < KR; ++t)
            if (t < k) row[t] = y[t];
    }
    // forward through R: entry e rotates columns (i, i+1), (c, s) as the
    // replay applies it
    const long ntot = act ? (long)rec[S.cntoff] : 0;
    const double* rs = rot + pt * 2 * rstride;
    const short* rx = ridx + pt * split_ridx_stride(kmax);
    long emax = ntot;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const long t = __shfl_xor(emax, o, 64);
        emax = t > emax ? t : emax;
    }
    // the stream is read kEvfChunk entries at a time, all loads issued before
    // the chunk's rotations run
    for (long e0 = 0; e0 < emax; e0 += kEvfChunk) {
        double cc[kEvfChunk], ss[kEvfChunk];
        int ii[kEvfChunk];
        evf_chunk(cc, ss, ii, rs, rx, e0, ntot);
#pragma unroll
        for (int u = 0; u < kEvfChunk; ++u) {
            const int i = ii[u];
            if (i >= 0) {
                const double a0 = row[i], a1 = row[i + 1];
                row[i] = fma(cc[u], a0, -ss[u] * a1);
                row[i + 1] = fma(ss[u], a0, cc[u] * a1);
            }
        }
    }
    __syncthreads();
    // u = (b^T V) / lambda in b's lane; truncation test on lambda
    double lmax = 0.0, lmin = 1.0e308;
    if (act)
        for (int t = 0; t < k; ++t) {
            const double lt = rec[S.lamoff + t];
            lmax = lt > lmax ? lt : lmax;
            lmin = lt < lmin ? lt : lmin;
            if (isb) row[t] = row[t] / lt;
        }
    __syncthreads();
    // X' w-bar = (x V) . u, then x V diag(U)
    double cw = 0.0;
    if (act && isx) {
        const double* u = sm + (size_t)(grp * G + nvmax) * ld;
        for (int t = 0; t < k; ++t) cw = fma(row[t], u[t], cw);
        for (int t = 0; t < k; ++t) row[t] = row[t] * sqrt((double)(k - 1) / rec[S.lamoff + t]);
    }
    __syncthreads();
    // backward through R^T: the entries in reverse, each rotation transposed
    for (long e0 = emax > 0 ? ((emax - 1) / kEvfChunk) * kEvfChunk : -1; e0 >= 0;
         e0 -= kEvfChunk) {
        double cc[kEvfChunk], ss[kEvfChunk];
        int ii[kEvfChunk];
        evf_chunk(cc, ss, ii, rs, rx, e0, ntot);
#pragma unroll
        for (int u = kEvfChunk - 1; u >= 0; --u) {
            const int i = ii[u];
            if (i >= 0) {
                const double a0 = row[i], a1 = row[i + 1];
                row[i] = fma(cc[u], a0, ss[u] * a1);
                row[i + 1] = fma(-ss[u], a0, cc[u] * a1);
            }
        }
    }
    // out = row Q^T (Q column-major at qoff: Q[m][t] at qoff + t * k + m, i.e.
    // Q^T row-major)
    if (coop) {
        __syncthreads();
        evf_coop_mv<8>(sm + (size_t)grp * G * ld, ld, G, k, rec + S.qoff, v, solve);
    }
    if (act) {
        double out[KR];
#pragma unroll
        for (int m = 0; m < KR; ++m) out[m] = 0.0;
        if (coop) {
#pragma unroll
            for (int m = 0; m < KR; ++m)
                if (m < k) out[m] = row[m];
        } else {
            const double* qc = rec + S.qoff;
            for (int t = 0; t < k; ++t) {
                const double zt = row[t];
                const double* qt = qc + (size_t)t * k;
#pragma unroll
                for (int m = 0; m < KR; ++m)
                    if (m < k) out[m] = fma(zt, qt[m], out[m]);
            }
        }
        const LekfJob& J = *Jp;
        if (isx) {
#pragma unroll
            for (int m = 0; m < KR; ++m)
                if (m < k) xdst[m] = upd ? mean + (cw + out[m]) : mean + (xsrc[m] - mean);
        } else {
            // w-bar, inflation, status
            if (J.transm)
#pragma unroll
                for (int i = 0; i < KR; ++i)
                    if (i < k) J.transm[(size_t)b * k + i] = out[i];
            if ((J.flags & LEKF_F_MINFL) && J.minfl > 0.0 && alpha < J.minfl) alpha = J.minfl;
            if (J.infl_out) {
                double tr = 0.0;
                for (int i = 0; i < k; ++i) tr += rec[(size_t)i * k + i];
                const double p1 = rec[(size_t)k * k + k], p3 = rec[(size_t)k * k + k + 1];
                J.infl_out[b] = (J.flags & LEKF_F_FIXED_INFL)
                                    ? J.infl_fixed
                                    : adaptive_alpha(alpha, p1, tr / (double)(k - 1), p3);
            }
            if (J.status) {
                int status = (((int)rec[S.flagoff]) & 1) ? LEKF_PT_NOCONV : 0;
                if (!(lmax > 0.0) || lmin < fabs(lmax) * 1.4901161193847656e-08)
                    status |= LEKF_PT_RANKDEF;
                J.status[b] = status | (pass ? LEKF_PT_FP64 : 0);
            }
        }
    }
}

// The points of a split launch whose rotation stream overflowed: the fused
// eigen kernel's work (from the Gram still at the start of the record), the
// others return at once.
template <int KR>
__global__ __launch_bounds__(64) void lekf_eig_fix_kernel(const LekfJob* __restrict__ jobs,
                                                           int njobs, long gp0,
                                                           double* __restrict__ ws, long gstride,
                                                           int nvmax, int kmax, int pass,
                                                           const int* __restrict__ rlist)
{
    extern __shared__ double smem[];
    long gp;
    if (!split_slot_point(blockIdx.x, gp0, (long)gridDim.x, pass, rlist, gp)) return;
    const LekfJob& J = jobs[lekf_job_of(jobs, njobs, gp)];
    const int b = (int)(gp - J.pt_begin + J.pt0);
    const int p = J.loc_cnt ? J.loc_cnt[b] : (int)(J.loc_off[b + 1] - J.loc_off[b]);
    double* rec = ws + (long)blockIdx.x * gstride;
    const SplitRec S = split_rec(kmax, nullptr);
    if (p <= 0 || split_other_pass(rec, S.flagoff, pass) ||
        !(((int)rec[S.flagoff]) & kSplitOverflow))
        return;
    eig_point<KR, 0>(J, b, rec, gstride, nvmax, 0L, smem, threadIdx.x);
    if (pass && threadIdx.x == 0 && J.status) J.status[b] |= LEKF_PT_FP64;
}

// LEKF_KSPLIT < k: one persistent wave per workgroup slot strides over the
// points; for each it forms the Gram blocks on MFMA into the slot's
// workspace record, then runs the eigen/assembly/apply steps in place.  The
// slot's record is reused from point to point, so the workspace is
// slots x record whatever the number of points, and one launch covers them
// all (points of very different cost balance out over the stride).
template <int MODE>
__global__ __launch_bounds__(64) void lekf_big_kernel(const LekfJob* __restrict__ jobs, int njobs,
                                                      long npt, double* __restrict__ ws,
                                                      long gstride, int nvmax, long rot_off,
                                                      int nblk)
{
    extern __shared__ double smem[];
    const int lane = threadIdx.x;
    double* const rec = ws + (size_t)blockIdx.x * gstride;
    for (long gp = blockIdx.x; gp < npt; gp += gridDim.x) {
        const LekfJob& J = jobs[lekf_job_of(jobs, njobs, gp)];
        const int b = (int)(gp - J.pt_begin + J.pt0);
        __syncthreads();  // the previous point is done with the record
        for (int t = 0; t < nblk; ++t) gram_blk_point<MODE>(J, b, t, rec, lane);
        __syncthreads();
        eig_point<0, 4>(J, b, rec, gstride, nvmax, rot_off, smem, lane);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host-side launchers (called from lekf_capi.hip)
// ---------------------------------------------------------------------------
int lekf_solve_nt(int k) { return (k + 1 + 15) / 16; }

// Kernel class of an ensemble size: KR * 8 + NT for k <= 32 (register
// Jacobi, KR = k rounded up to a multiple of 4); NT for k <= LEKF_KSPLIT;
// kClassBig above (block Gram + eigensolver in global memory).
constexpr int kClassBig = 1000;

int lekf_solve_class(int k)
{
    if (k <= 32) {
        const int kr = (k + 3) & ~3;
        const int nt = kr <= 12 ? 1 : (kr <= 28 ? 2 : 3);
        return kr * 8 + nt;
    }
    if (k > LEKF_KSPLIT) return kClassBig;
    return lekf_solve_nt(k);
}

// points per wave of the fused kernel for Jacobi rows KR: lane groups of KR,
// at most 4 points (the LDS holds a region per point)
static constexpr int fused_ppw(int kr) { return 64 / kr < 4 ? 64 / kr : 4; }

size_t lekf_solve_lds_bytes(int k, int nvmax)
{
    if (k <= 32) return solve_lds_doubles(k, nvmax, fused_ppw((k + 3) & ~3)) * sizeof(double);
    if (k > LEKF_KSPLIT) return eig_lds_doubles_big(k, nvmax) * sizeof(double);
    return eig_lds_doubles(k, nvmax) * sizeof(double);
}

// register eigensolver variant of an ensemble size: k rounded up to a multiple
// of 4 for k <= 64, 0 (LDS eigensolver) above
static int eig_kr(int kmax) { return kmax <= 64 ? (kmax + 3) & ~3 : 0; }

static bool split_path(int kmax)
{
    return kmax > 32 && kmax <= LEKF_KSPLIT && g_lekf_tune.eig_split && eig_kr(kmax) > 0;
}

// bytes per point of a sub-batch: the record, plus the rotation stream on the
// split path
static size_t ws_rec_bytes(int kmax)
{
    if (kmax > LEKF_KSPLIT) return (size_t)big_stride(kmax) * sizeof(double);
    if (split_path(kmax))  // record, rotation stream, its columns, a redo-list entry
        return (size_t)(split_stride(kmax) + 2 * split_rcap(kmax) + split_ridx_doubles(kmax) + 1) *
               sizeof(double);
    return (size_t)gram_stride(kmax) * sizeof(double);
}

size_t lekf_solve_ws_bytes(long npt, int kmax)
{
    if (kmax <= 32 || npt <= 0) return 0;
    const bool big = kmax > LEKF_KSPLIT;
    const size_t rec = ws_rec_bytes(kmax);
    size_t per = g_lekf_tune.gram_ws_bytes / rec;
    if (per < 1) per = 1;
    // large-k records are MBs each, and the split path's QL kernel takes 64
    // points per wave: keep at least 1024 points per launch
    if ((big || split_path(kmax)) && per < 1024) per = 1024;
    if ((size_t)npt < per) per = (size_t)npt;
    return per * rec;
}

template <int NT, int KR, int MODE>
static hipError_t launch_fused(const LekfJob* djobs, int njobs, long npt, int kmax, int nvmax,
                               hipStream_t st)
{
    constexpr int PPW = fused_ppw(KR);
    if (kmax > KR) return hipErrorInvalidValue;
    const size_t lds = solve_lds_doubles(kmax, nvmax, PPW) * sizeof(double);
    const long nwg = (npt + PPW - 1) / PPW;
    KTime kt("solve", st);
    hipLaunchKernelGGL((lekf_solve_kernel<NT, KR, MODE, PPW>), dim3((unsigned)nwg), dim3(LEKF_WAVE),
                       lds, st, djobs, njobs, npt, kmax, nvmax);
    return hipGetLastError();
}

static hipError_t launch_eig(int kr, long nb, size_t elds, hipStream_t st, const LekfJob* djobs,
                             int njobs, long g0, double* ws, long gs, int nvmax)
{
    KTime kt("eig", st);
    switch (kr) {
#define LEKF_EIG_CASE(K)                                                                       \
    case K:                                                                                    \
        hipLaunchKernelGGL(lekf_eig_kernel<K>, dim3((unsigned)nb), dim3(LEKF_WAVE), elds, st, \
                           djobs, njobs, g0, ws, gs, nvmax);                                  \
        break;
        LEKF_EIG_CASE(36)
        LEKF_EIG_CASE(40)
        LEKF_EIG_CASE(44)
        LEKF_EIG_CASE(48)
        LEKF_EIG_CASE(52)
        LEKF_EIG_CASE(56)
        LEKF_EIG_CASE(60)
        LEKF_EIG_CASE(64)
        LEKF_EIG_CASE(0)
#undef LEKF_EIG_CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

static hipError_t launch_eig_split(int kr, long nb, size_t elds, hipStream_t st,
                                   const LekfJob* djobs, int njobs, long g0, double* ws, long gs,
                                   int nvmax, int kmax, double* rot, short* ridx, int pass,
                                   double kthr, int* rlist, int zero, long lcap)
{
    const unsigned nwq = (unsigned)((nb + 63) / 64);
    const int evp = 64 / (nvmax + 1);  // points per wave of the eigenvector-free kernel
    const unsigned nwe = (unsigned)((nb + evp - 1) / evp);
    const size_t flds = 64 * (size_t)(kmax | 1) * sizeof(double);
    const size_t qlds = 2 * (size_t)kr * 64 * sizeof(double);
    const size_t tlds = tri_lds_doubles(kmax) * sizeof(double);
    switch (kr) {
#define LEKF_SPLIT_CASE(K)                                                                       \
    case K: {                                                                                    \
        {                                                                                        \
            KTime kt(pass ? "redo_tri" : "eig_tri", st);                                         \
            hipLaunchKernelGGL(lekf_tri_kernel, dim3((unsigned)nb), dim3(LEKF_WAVE), tlds, st,    \
                               djobs, njobs, g0, ws, gs, kmax, pass, rlist, zero);              \
        }                                                                                        \
        {                                                                                        \
            KTime kt(pass ? "redo_ql" : "eig_ql", st);                                           \
            hipLaunchKernelGGL(lekf_ql_kernel<K>, dim3(nwq), dim3(LEKF_WAVE), qlds, st, djobs,    \
                               njobs, g0, nb, ws, gs, rot, ridx, kmax,                           \
                               g_lekf_tune.split_rcap, pass, kthr, rlist, lcap);                \
        }                                                                                        \
        KTime kt(pass ? "redo_apply" : "eig_apply", st);                                         \
        hipLaunchKernelGGL(lekf_eig3_kernel<K>, dim3((unsigned)nb), dim3(LEKF_WAVE), elds, st,    \
                           djobs, njobs, g0, ws, gs, nvmax, kmax, rot, pass, rlist);            \
        hipLaunchKernelGGL(lekf_evf_kernel<K>, dim3(nwe), dim3(LEKF_WAVE), flds, st, djobs,      \
                           njobs, g0, nb, ws, gs, rot, ridx, nvmax, kmax, pass, rlist);         \
        hipLaunchKernelGGL(lekf_eig_fix_kernel<K>, dim3((unsigned)nb), dim3(LEKF_WAVE), elds, st, \
                           djobs, njobs, g0, ws, gs, nvmax, kmax, pass, rlist);                 \
        break;                                                                                   \
    }
        LEKF_SPLIT_CASE(36)
        LEKF_SPLIT_CASE(40)
        LEKF_SPLIT_CASE(44)
        LEKF_SPLIT_CASE(48)
        LEKF_SPLIT_CASE(52)
        LEKF_SPLIT_CASE(56)
        LEKF_SPLIT_CASE(60)
        LEKF_SPLIT_CASE(64)
#undef LEKF_SPLIT_CASE
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int NT, int MODE>
static hipError_t launch_gram(int W, long nb, size_t glds, hipStream_t st, const LekfJob* djobs,
                              int njobs, long g0, double* ws, long gs, const int* rlist)
{
    KTime kt(rlist ? "redo_gram" : "gram", st);
    if (W == 2)
        hipLaunchKernelGGL((lekf_gram_kernel<NT, MODE, 2>), dim3((unsigned)nb), dim3(128), glds, st,
                           djobs, njobs, g0, ws, gs, rlist);
    else if (W == 8)
        hipLaunchKernelGGL((lekf_gram_kernel<NT, MODE, 8>), dim3((unsigned)nb), dim3(512), glds, st,
                           djobs, njobs, g0, ws, gs, rlist);
    else
        hipLaunchKernelGGL((lekf_gram_kernel<NT, MODE, 4>), dim3((unsigned)nb), dim3(256), glds, st,
                           djobs, njobs, g0, ws, gs, rlist);
    return hipGetLastError();
}

template <int NT, int MODE>
static hipError_t launch_split(const LekfJob* djobs, int njobs, long npt, int kmax, int nvmax,
                               void* ws, size_t ws_bytes, hipStream_t st)
{
    const bool s3 = split_path(kmax);
    const long gs = s3 ? split_stride(kmax) : gram_stride(kmax);
    const long per = (long)(ws_bytes / ws_rec_bytes(kmax));
    if (per < 1 || !ws) return hipErrorInvalidValue;
    double* rot = (double*)ws + (size_t)per * gs;  // the rotation streams after the records
    short* ridx = (short*)(rot + (size_t)per * 2 * split_rcap(kmax));  // then their columns
    int* rlist = (int*)(ridx + (size_t)per * split_ridx_stride(kmax));  // then the redo list
    const int W = g_lekf_tune.gram_waves;
    const size_t glds = ((size_t)kmax * (kmax + 1) + kmax + 4 + W) * sizeof(double);
    const size_t elds = eig_lds_doubles(kmax, nvmax) * sizeof(double);
    // mixed forms on the split eigensolver: the QL kernel lists the points whose
    // C is worse conditioned than kthr (up to one sub-batch of them per launch),
    // and after the last sub-batch one FP64 pass solves the listed points again
    // (one set of kernel latencies per launch, whatever the sub-batch count).
    // The FP32 Gram rounds about 16 times finer than split-BF16x3, so its
    // threshold is 16 times the value set.
    const double kthr = (MODE != 0 && s3 && npt <= 0x7fffffffL)
                            ? (MODE == 1 ? 16.0 : 1.0) * (double)g_lekf_tune.mixed_kappa
                            : 0.0;
    for (long g0 = 0; g0 < npt; g0 += per) {
        const long nb = per < npt - g0 ? per : npt - g0;
        hipError_t e =
            launch_gram<NT, MODE>(W, nb, glds, st, djobs, njobs, g0, (double*)ws, gs, nullptr);
        if (e != hipSuccess) return e;
        if (s3)
            e = launch_eig_split(eig_kr(kmax), nb, elds, st, djobs, njobs, g0, (double*)ws, gs,
                                 nvmax, kmax, rot, ridx, 0, kthr, rlist, g0 == 0, per);
        else
            e = launch_eig(eig_kr(kmax), nb, elds, st, djobs, njobs, g0, (double*)ws, gs, nvmax);
        if (e != hipSuccess) return e;
    }
    if (kthr > 0.0) {
        // grids of one sub-batch; the slots past the list's length return at once
        const long nl = per < npt ? per : npt;
        hipError_t e = launch_gram<NT, 0>(W, nl, glds, st, djobs, njobs, 0, (double*)ws, gs, rlist);
        if (e == hipSuccess)
            e = launch_eig_split(eig_kr(kmax), nl, elds, st, djobs, njobs, 0, (double*)ws, gs, nvmax,
                                 kmax, rot, ridx, 1, 0.0, rlist, 0, per);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// LEKF_KSPLIT < k <= LEKF_KMAX: block Gram (one wave per 64 x 64 block of a
// point), then the eigen/assembly kernel working in place on each point's
// workspace record
template <int MODE>
static hipError_t launch_big(const LekfJob* djobs, int njobs, long npt, int kmax, int nvmax,
                             void* ws, size_t ws_bytes, hipStream_t st)
{
    const long gs = big_stride(kmax);
    const long slots = (long)(ws_bytes / ((size_t)gs * sizeof(double)));
    if (slots < 1 || !ws) return hipErrorInvalidValue;
    const size_t elds = eig_lds_doubles_big(kmax, nvmax) * sizeof(double);
    // as many workgroups as can be resident at once (LDS bounds the count)
    int dev = 0, ncu = 256, per_cu = 1;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, lekf_big_kernel<MODE>, LEKF_WAVE, elds);
    long nwg = (long)(per_cu > 0 ? per_cu : 1) * (ncu > 0 ? ncu : 256);
    if (nwg > slots) nwg = slots;
    if (nwg > npt) nwg = npt;
    KTime kt("big", st);
    hipLaunchKernelGGL((lekf_big_kernel<MODE>), dim3((unsigned)nwg), dim3(LEKF_WAVE), elds, st, djobs,
                       njobs, npt, (double*)ws, gs, nvmax, (long)gram_stride(kmax), gram_nblk(kmax));
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_mode(int cls, const LekfJob* djobs, int njobs, long npt, int kmax,
                              int nvmax, void* ws, size_t ws_bytes, hipStream_t st)
{
    switch (cls) {
    case 4 * 8 + 1: return launch_fused<1, 4, MODE>(djobs, njobs, npt, kmax, nvmax, st);
    case 8 * 8 + 1: return launch_fused<1, 8, MODE>(djobs, njobs, npt, kmax, nvmax, st);
    case 12 * 8 + 1: return launch_fused<1, 12, MODE>(djobs, njobs, npt, kmax, nvmax, st);
    case 16 * 8 + 2: return launch_fused<2, 16, MODE>(djobs, njobs, npt, kmax, nvmax, st);
    case 20 * 8 + 2: return launch_fused<2, 20, MODE>(djobs, njobs, npt, kmax, nvmax, st);
    case 24 * 8 + 2: return launch_fused<2, 24, MODE>(djobs, njobs, npt, kmax, nvmax, st);
    case 28 * 8 + 2: return launch_fused<2, 28, MODE>(djobs, njobs, npt, kmax, nvmax, st);
    case 32 * 8 + 3: return launch_fused<3, 32, MODE>(djobs, njobs, npt, kmax, nvmax, st);
    case 3: return launch_split<3, MODE>(djobs, njobs, npt, kmax, nvmax, ws, ws_bytes, st);
    case 4: return launch_split<4, MODE>(djobs, njobs, npt, kmax, nvmax, ws, ws_bytes, st);
    case 5: return launch_split<5, MODE>(djobs, njobs, npt, kmax, nvmax, ws, ws_bytes, st);
    case 6: return launch_split<6, MODE>(djobs, njobs, npt, kmax, nvmax, ws, ws_bytes, st);
    case kClassBig: return launch_big<MODE>(djobs, njobs, npt, kmax, nvmax, ws, ws_bytes, st);
    default: return hipErrorInvalidValue;
    }
}

// Launch the solve over npt flat points of njobs jobs (device array djobs).
// Every job of a launch has the same kernel class; kmax (the largest k) and
// nvmax (the most state rows a point applies to) size the LDS; ws is the
// Gram workspace of lekf_solve_ws_bytes(npt, kmax) bytes (k > 32).
hipError_t lekf_solve_launch(const LekfJob* djobs, int njobs, long npt, int kmax, int nvmax,
                             int mode, void* ws, size_t ws_bytes, hipStream_t st)
{
    if (npt <= 0) return hipSuccess;
    if (nvmax > LEKF_NVMAX) return hipErrorInvalidValue;
    const int cls = lekf_solve_class(kmax);
    if (mode == 0) return launch_mode<0>(cls, djobs, njobs, npt, kmax, nvmax, ws, ws_bytes, st);
    if (mode == 1) return launch_mode<1>(cls, djobs, njobs, npt, kmax, nvmax, ws, ws_bytes, st);
    return launch_mode<2>(cls, djobs, njobs, npt, kmax, nvmax, ws, ws_bytes, st);
}
